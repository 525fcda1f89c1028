// oneint.hip -- one-electron overlap, kinetic and nuclear-attraction
// matrices, one thread per shell pair, one kernel instance per pair class.
//
// Per primitive pair (jax_qc/integrals/obara_saika.py:70-222):
//   S_ij^x = sqrt(pi/p) E_ij^0,  S = K S^x S^y S^z
//   T_ij^x = 1/2 [4ab S_{i+1,j+1} + ij S_{i-1,j-1} - 2aj S_{i+1,j-1} - 2bi S_{i-1,j+1}]
//   T = K (T^x S^y S^z + S^x T^y S^z + S^x S^y T^z)
//   V = -sum_C Z_C (2 pi/p) K sum_tuv E^x_t E^y_u E^z_v R_tuv(p, P-C)
//   dipole (which = 3, 4, 5 for x, y, z about the origin O):
//   M_ij^x = sqrt(pi/p) (E_ij^1 + (P_x - O_x) E_ij^0),  D_x = K M^x S^y S^z
// (the first Hermite moment: int x_O Lambda_t = sqrt(pi/p) [t = 1] +
// X_PO sqrt(pi/p) [t = 0]).
// K already carries the contraction coefficients.  The Cartesian block goes to
// the spherical basis (overlap.py:58-64) and is written to both triangles.
#include "qcmd_device.h"
#include "qcmd_internal.h"
#include "qcmd_tables.h"
#include "qcmd_types.h"

namespace qcmd {

template <int LA, int LB>
__global__ __launch_bounds__(128) void one_electron_kernel(
    int which, const double* __restrict__ pp, const int* __restrict__ sp_prim_off,
    const int* __restrict__ sp_nprim, const int* __restrict__ sp_ao, const int* __restrict__ list,
    int npairs, int natom, const double* __restrict__ axyz, const double* __restrict__ az,
    int64_t nbf, double* __restrict__ M, const double* __restrict__ boys,
    const int* __restrict__ pair_atoms, const int64_t* __restrict__ pair_base) {
  constexpr int NA = ncart(LA), NB = ncart(LB), L = LA + LB;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= npairs) return;
  const int sp = list[idx];
  // many molecules in one launch: the pair's own nuclei and output block
  const int c0 = pair_atoms ? pair_atoms[2 * sp] : 0;
  const int c1 = pair_atoms ? c0 + pair_atoms[2 * sp + 1] : natom;
  if (pair_base) M += pair_base[sp];
  const int p0 = sp_prim_off[sp], np = sp_nprim[sp];
  double acc[NA][NB];
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = 0.0;

  for (int ip = 0; ip < np; ++ip) {
    const double* P = pp + (int64_t)(p0 + ip) * PP_STRIDE;
    const double p = P[PP_P], K = P[PP_K], al = P[PP_ALPHA], be = p - al;
    const double oo2p = 0.5 / p;
    if (which == 2) {
      double Ex[LA + 1][LB + 1][L + 1], Ey[LA + 1][LB + 1][L + 1], Ez[LA + 1][LB + 1][L + 1];
      hermite_e<LA, LB, double>(Ex, P[PP_PAX], P[PP_PBX], oo2p);
      hermite_e<LA, LB, double>(Ey, P[PP_PAX + 1], P[PP_PBX + 1], oo2p);
      hermite_e<LA, LB, double>(Ez, P[PP_PAX + 2], P[PP_PBX + 2], oo2p);
      for (int c = c0; c < c1; ++c) {
        const double X = P[PP_PX] - axyz[3 * c], Y = P[PP_PX + 1] - axyz[3 * c + 1],
                     Z = P[PP_PX + 2] - axyz[3 * c + 2];
        double F[L + 1];
        boys_fn<L, double>(p * (X * X + Y * Y + Z * Z), boys, F);
        double R[L + 1][L + 1][L + 1];
        hermite_r<L, double>(R, F, -2.0 * p, X, Y, Z);
        const double pref = -az[c] * (6.283185307179586477 / p) * K;
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            const int sx = CX[LA][a] + CX[LB][b], sy = CY[LA][a] + CY[LB][b],
                      sz = CZ[LA][a] + CZ[LB][b];
            double v = 0.0;
#pragma unroll
            for (int t = 0; t <= L; ++t) {
              if (t > sx) continue;
#pragma unroll
              for (int u = 0; u <= L; ++u) {
                if (u > sy) continue;
#pragma unroll
                for (int w = 0; w <= L; ++w) {
                  if (w > sz) continue;
                  v += Ex[CX[LA][a]][CX[LB][b]][t] * Ey[CY[LA][a]][CY[LB][b]][u] *
                       Ez[CZ[LA][a]][CZ[LB][b]][w] * R[t][u][w];
                }
              }
            }
            acc[a][b] += pref * v;
          }
      }
    } else {
      // 1-D overlaps up to (LA+1, LB+1) for the kinetic shifts.
      double Ex[LA + 2][LB + 2][L + 3], Ey[LA + 2][LB + 2][L + 3], Ez[LA + 2][LB + 2][L + 3];
      hermite_e<LA + 1, LB + 1, double>(Ex, P[PP_PAX], P[PP_PBX], oo2p);
      hermite_e<LA + 1, LB + 1, double>(Ey, P[PP_PAX + 1], P[PP_PBX + 1], oo2p);
      hermite_e<LA + 1, LB + 1, double>(Ez, P[PP_PAX + 2], P[PP_PBX + 2], oo2p);
      const double f = sqrt(3.14159265358979323846 / p);
      double Sx[LA + 2][LB + 2], Sy[LA + 2][LB + 2], Sz[LA + 2][LB + 2];
#pragma unroll
      for (int i = 0; i <= LA + 1; ++i)
#pragma unroll
        for (int j = 0; j <= LB + 1; ++j) {
          Sx[i][j] = f * Ex[i][j][0];
          Sy[i][j] = f * Ey[i][j][0];
          Sz[i][j] = f * Ez[i][j][0];
        }
      double Tx[LA + 1][LB + 1], Ty[LA + 1][LB + 1], Tz[LA + 1][LB + 1];
      // dipole: first moment about the origin (axyz) in the field direction
      // (three branches keep every table index compile-time)
      if (which == 3) {
        const double xo = P[PP_PX] - axyz[0];
#pragma unroll
        for (int i = 0; i <= LA; ++i)
#pragma unroll
          for (int j = 0; j <= LB; ++j) Sx[i][j] = f * (Ex[i][j][1] + xo * Ex[i][j][0]);
      } else if (which == 4) {
        const double yo = P[PP_PX + 1] - axyz[1];
#pragma unroll
        for (int i = 0; i <= LA; ++i)
#pragma unroll
          for (int j = 0; j <= LB; ++j) Sy[i][j] = f * (Ey[i][j][1] + yo * Ey[i][j][0]);
      } else if (which == 5) {
        const double zo = P[PP_PX + 2] - axyz[2];
#pragma unroll
        for (int i = 0; i <= LA; ++i)
#pragma unroll
          for (int j = 0; j <= LB; ++j) Sz[i][j] = f * (Ez[i][j][1] + zo * Ez[i][j][0]);
      }
      if (which == 1) {
#pragma unroll
        for (int i = 0; i <= LA; ++i)
#pragma unroll
          for (int j = 0; j <= LB; ++j) {
            double tx = 4.0 * al * be * Sx[i + 1][j + 1], ty = 4.0 * al * be * Sy[i + 1][j + 1],
                   tz = 4.0 * al * be * Sz[i + 1][j + 1];
            if (i > 0 && j > 0) {
              tx += i * j * Sx[i - 1][j - 1];
              ty += i * j * Sy[i - 1][j - 1];
              tz += i * j * Sz[i - 1][j - 1];
            }
            if (j > 0) {
              tx -= 2.0 * al * j * Sx[i + 1][j - 1];
              ty -= 2.0 * al * j * Sy[i + 1][j - 1];
              tz -= 2.0 * al * j * Sz[i + 1][j - 1];
            }
            if (i > 0) {
              tx -= 2.0 * be * i * Sx[i - 1][j + 1];
              ty -= 2.0 * be * i * Sy[i - 1][j + 1];
              tz -= 2.0 * be * i * Sz[i - 1][j + 1];
            }
            Tx[i][j] = 0.5 * tx;
            Ty[i][j] = 0.5 * ty;
            Tz[i][j] = 0.5 * tz;
          }
      }
#pragma unroll
      for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const int ax = CX[LA][a], ay = CY[LA][a], az_ = CZ[LA][a];
          const int bx = CX[LB][b], by = CY[LB][b], bz = CZ[LB][b];
          double v;
          if (which == 0 || which >= 3)
            v = Sx[ax][bx] * Sy[ay][by] * Sz[az_][bz];
          else
            v = Tx[ax][bx] * Sy[ay][by] * Sz[az_][bz] + Sx[ax][bx] * Ty[ay][by] * Sz[az_][bz] +
                Sx[ax][bx] * Sy[ay][by] * Tz[az_][bz];
          acc[a][b] += K * v;
        }
    }
  }

  constexpr int SA = nsph(LA), SB = nsph(LB);
  const int64_t ia = sp_ao[2 * sp], ib = sp_ao[2 * sp + 1];
#pragma unroll
  for (int sa = 0; sa < SA; ++sa)
#pragma unroll
    for (int sb = 0; sb < SB; ++sb) {
      double v = 0.0;
#pragma unroll
      for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b)
          if (SPH[LA][sa][a] != 0.0 && SPH[LB][sb][b] != 0.0)
            v += SPH[LA][sa][a] * SPH[LB][sb][b] * acc[a][b];
      M[(ia + sa) * nbf + ib + sb] = v;
      M[(ib + sb) * nbf + ia + sa] = v;
    }
}

}  // namespace qcmd

using namespace qcmd;

static int one_electron(int which, int cab, const double* pairs64, const int* sp_prim_off,
                        const int* sp_nprim, const int* sp_ao, const int* pair_list, int npairs,
                        int natom, const double* atom_xyz, const double* atom_z, int64_t nbf,
                        double* M, const int* pair_atoms, const int64_t* pair_base, void* stream) {
  if (which < 0 || which > 5) return set_error("qc_one_electron: which=%d", which);
  if (which >= 3 && !atom_xyz) return set_error("qc_one_electron: dipole needs an origin");
  if (npairs <= 0) return 0;
  const double* boys = boys_table64();
  if (!boys) return set_error("qc_one_electron: qc_init not called on the current device");
  hipStream_t s = (hipStream_t)stream;
  dim3 grid((npairs + 127) / 128), block(128);
#define QC_ONE(A, B)                                                                          \
  case (A) * ((A) + 1) / 2 + (B):                                                             \
    hipLaunchKernelGGL((one_electron_kernel<A, B>), grid, block, 0, s, which, pairs64,        \
                       sp_prim_off, sp_nprim, sp_ao, pair_list, npairs, natom, atom_xyz, atom_z, \
                       nbf, M, boys, pair_atoms, pair_base);                                  \
    break;
  switch (cab) {
    QC_ONE(0, 0) QC_ONE(1, 0) QC_ONE(1, 1) QC_ONE(2, 0) QC_ONE(2, 1) QC_ONE(2, 2)
    default:
      return set_error("qc_one_electron: pair class %d", cab);
  }
#undef QC_ONE
  return check(hipGetLastError(), "qc_one_electron launch");
}

extern "C" int qc_one_electron(int which, int cab, const double* pairs64, const int* sp_prim_off,
                               const int* sp_nprim, const int* sp_ao, const int* pair_list,
                               int npairs, int natom, const double* atom_xyz, const double* atom_z,
                               int64_t nbf, double* M, void* stream) {
  return one_electron(which, cab, pairs64, sp_prim_off, sp_nprim, sp_ao, pair_list, npairs, natom,
                      atom_xyz, atom_z, nbf, M, nullptr, nullptr, stream);
}

extern "C" int qc_one_electron_batch(int which, int cab, const double* pairs64,
                                     const int* sp_prim_off, const int* sp_nprim, const int* sp_ao,
                                     const int* pair_list, int npairs, const int* pair_atoms,
                                     const double* atom_xyz, const double* atom_z, int64_t nmax,
                                     const int64_t* pair_base, double* M, void* stream) {
  if (!pair_base || (which == 2 && !pair_atoms))
    return set_error("qc_one_electron_batch: null pair_base or pair_atoms");
  return one_electron(which, cab, pairs64, sp_prim_off, sp_nprim, sp_ao, pair_list, npairs, 0,
                      atom_xyz, atom_z, nmax, M, pair_atoms, pair_base, stream);
}
