// eri_impl.h -- shell-quartet electron-repulsion integrals by McMurchie-Davidson
// on gfx950, one thread per contracted quartet, one kernel instance per
// angular-momentum class (la lb | lc ld) with la >= lb, lc >= ld and the bra
// pair class >= the ket pair class.
//
// Per primitive quartet (jax_qc/integrals/obara_saika.py:227-332):
//   (ab|cd) = 2 pi^{5/2} / (p q sqrt(p+q)) K_AB K_CD
//             sum_{tuv} E^{ab}_{tuv} sum_{t'u'v'} (-1)^{t'+u'+v'} E^{cd}_{t'u'v'}
//             R_{t+t',u+u',v+v'}(p q/(p+q), P-Q)
// evaluated as: for each ket Cartesian pair cd, h_{tuv} = ket-contracted
// Hermite integrals (t+u+v <= la+lb); then each bra Cartesian pair ab takes
// its sparse E^{ab} dot h.  Contraction over primitives accumulates in
// registers; the Cartesian block is taken to the spherical basis once per
// contracted quartet (jax_qc/integrals/eri.py:61-66) and written out as a
// dense tensor (8-fold scatter, eri.py:108-138), into the packed supermatrix,
// or reduced to a Schwarz bound (screening.py:32-58).
#pragma once
#include "qcmd_device.h"
#include "qcmd_tables.h"
#include "qcmd_types.h"

namespace qcmd {

template <typename Real>
__device__ inline const Real* pair_table(const EriArgs& a);
template <>
__device__ inline const double* pair_table<double>(const EriArgs& a) { return a.pairs64; }
template <>
__device__ inline const float* pair_table<float>(const EriArgs& a) { return a.pairs32; }
template <typename Real>
__device__ inline const Real* boys_table(const EriArgs& a);
template <>
__device__ inline const double* boys_table<double>(const EriArgs& a) { return a.boys64; }
template <>
__device__ inline const float* boys_table<float>(const EriArgs& a) { return a.boys32; }

__device__ inline int64_t sm_pos(int64_t p, int64_t q) {  // p >= q
  int64_t R = p >> SM_LOGB, C = q >> SM_LOGB;
  return ((R * (R + 1) / 2 + C) << (2 * SM_LOGB)) + ((p & (SM_B - 1)) << SM_LOGB) + (q & (SM_B - 1));
}

__device__ inline void store_super(double* out, int64_t mu, int64_t nu, int64_t la, int64_t si,
                                   double v) {
  int64_t p = mu >= nu ? mu * (mu + 1) / 2 + nu : nu * (nu + 1) / 2 + mu;
  int64_t q = la >= si ? la * (la + 1) / 2 + si : si * (si + 1) / 2 + la;
  if (p < q) {
    int64_t t = p;
    p = q;
    q = t;
  }
  out[sm_pos(p, q)] = v;
  // diagonal tiles hold both triangles (sm_pos only needs p and q in the
  // same tile for the swapped in-tile offsets)
  if ((p >> SM_LOGB) == (q >> SM_LOGB)) out[sm_pos(q, p)] = v;
}

// Locate the quartet handled by global index gid: bra = last b with
// offsets[b] <= gid, ket position = kstart[b] + gid - offsets[b].
__device__ inline void locate(const EriArgs& a, int64_t gid, int& bp, int& kp) {
  int lo = 0, hi = a.nbra - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (a.offsets[mid] <= gid)
      lo = mid;
    else
      hi = mid - 1;
  }
  bp = a.bra_list[lo];
  kp = a.ket_list[a.kstart[lo] + (int)(gid - a.offsets[lo])];
}

// Ket Cartesian pairs cd in [CD0, CD1) only: for kets up to (pp) the ket's
// spherical transform is the identity, so a launch per range computes and
// writes exactly its own integrals (not used in Schwarz mode, which needs the
// whole diagonal block).  Likewise bra Cartesians b in [B0, B1) only, for a
// p (or s) shell b.
template <int LA, int LB, int LC, int LD, typename Real, int CD0, int CD1, int B0 = 0,
          int B1 = ncart(LB)>
__global__ __launch_bounds__(128) void eri_class_kernel(EriArgs args) {
  static_assert((B0 == 0 && B1 == ncart(LB)) || LB <= 1, "bra ranges need an identity b transform");
  constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD;
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  const int64_t gid = flat_block() * blockDim.x + threadIdx.x;
  if (gid >= args.nquartets) return;
  int bp, kp;
  locate(args, gid, bp, kp);

  const Real* __restrict__ pp = pair_table<Real>(args);
  const int pb0 = args.sp_prim_off[bp], npb = args.sp_nprim[bp];
  const int pk0 = args.sp_prim_off[kp], npk = args.sp_nprim[kp];

  Real acc[NA * NB][NC * ND];
#pragma unroll
  for (int i = 0; i < NA * NB; ++i)
#pragma unroll
    for (int j = 0; j < NC * ND; ++j) acc[i][j] = (Real)0;

  // Ket primitive pairs outermost: the (smaller) ket E tables stay live
  // while the bra tables are rebuilt per primitive quartet, which is cheaper
  // than holding them across the inner loop.
  for (int ik = 0; ik < npk; ++ik) {
    const Real* Q = pp + (int64_t)(pk0 + ik) * PP_STRIDE;
    const Real q = Q[PP_P], Qx = Q[PP_PX], Qy = Q[PP_PX + 1], Qz = Q[PP_PX + 2];
    const Real Kcd = Q[PP_K];
    const Real oo2q = (Real)0.5 / q;
    Real Ekx[LC + 1][LD + 1][LCD + 1], Eky[LC + 1][LD + 1][LCD + 1], Ekz[LC + 1][LD + 1][LCD + 1];
    hermite_e<LC, LD, Real>(Ekx, Q[PP_PAX], Q[PP_PBX], oo2q);
    hermite_e<LC, LD, Real>(Eky, Q[PP_PAX + 1], Q[PP_PBX + 1], oo2q);
    hermite_e<LC, LD, Real>(Ekz, Q[PP_PAX + 2], Q[PP_PBX + 2], oo2q);
    for (int ib = 0; ib < npb; ++ib) {
      const Real* P = pp + (int64_t)(pb0 + ib) * PP_STRIDE;
      const Real p = P[PP_P];
      const Real oo2p = (Real)0.5 / p;
      Real Ebx[LA + 1][LB + 1][LAB + 1], Eby[LA + 1][LB + 1][LAB + 1], Ebz[LA + 1][LB + 1][LAB + 1];
      hermite_e<LA, LB, Real>(Ebx, P[PP_PAX], P[PP_PBX], oo2p);
      hermite_e<LA, LB, Real>(Eby, P[PP_PAX + 1], P[PP_PBX + 1], oo2p);
      hermite_e<LA, LB, Real>(Ebz, P[PP_PAX + 2], P[PP_PBX + 2], oo2p);
      const Real pq = p * q, ps = p + q;
      const Real alpha = pq / ps;
      const Real X = P[PP_PX] - Qx, Y = P[PP_PX + 1] - Qy, Z = P[PP_PX + 2] - Qz;
      const Real T = alpha * (X * X + Y * Y + Z * Z);
      Real F[L + 1];
      boys_eval<L, Real>(T, args.boys64, F);
      Real R[L + 1][L + 1][L + 1];
      hermite_r<L, Real>(R, F, (Real)-2 * alpha, X, Y, Z);
      const Real pref = (Real)34.986836655249725693 / (pq * dev_sqrt(ps)) * P[PP_K] * Kcd;

#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int d = 0; d < ND; ++d) {
          if (c * ND + d < CD0 || c * ND + d >= CD1) continue;
          const int cx = CX[LC][c], cy = CY[LC][c], cz = CZ[LC][c];
          const int dx = CX[LD][d], dy = CY[LD][d], dz = CZ[LD][d];
          Real h[LAB + 1][LAB + 1][LAB + 1];
#pragma unroll
          for (int t = 0; t <= LAB; ++t)
#pragma unroll
            for (int u = 0; u <= LAB - t; ++u)
#pragma unroll
              for (int v = 0; v <= LAB - t - u; ++v) {
                Real s = (Real)0;
#pragma unroll
                for (int tt = 0; tt <= LCD; ++tt) {
                  if (tt > cx + dx) continue;
#pragma unroll
                  for (int uu = 0; uu <= LCD; ++uu) {
                    if (uu > cy + dy) continue;
#pragma unroll
                    for (int vv = 0; vv <= LCD; ++vv) {
                      if (vv > cz + dz) continue;
                      const Real e = Ekx[cx][dx][tt] * Eky[cy][dy][uu] * Ekz[cz][dz][vv];
                      if ((tt + uu + vv) & 1)
                        s -= e * R[t + tt][u + uu][v + vv];
                      else
                        s += e * R[t + tt][u + uu][v + vv];
                    }
                  }
                }
                h[t][u][v] = s;
              }
#pragma unroll
          for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int b = 0; b < NB; ++b) {
              if (b < B0 || b >= B1) continue;
              const int ax = CX[LA][a], ay = CY[LA][a], az = CZ[LA][a];
              const int bx = CX[LB][b], by = CY[LB][b], bz = CZ[LB][b];
              Real val = (Real)0;
#pragma unroll
              for (int t = 0; t <= LAB; ++t) {
                if (t > ax + bx) continue;
#pragma unroll
                for (int u = 0; u <= LAB; ++u) {
                  if (u > ay + by) continue;
#pragma unroll
                  for (int v = 0; v <= LAB; ++v) {
                    if (v > az + bz) continue;
                    val += Ebx[ax][bx][t] * Eby[ay][by][u] * Ebz[az][bz][v] * h[t][u][v];
                  }
                }
              }
              acc[a * NB + b][c * ND + d] += pref * val;
            }
        }
    }
  }

  if (args.mode == ERI_SCHWARZ) {
    // max over the Cartesian diagonal (ab|ab) (screening.py:47-58); only
    // launched on (P|P) quartets, so LA == LC and LB == LD.
    double mx = 0.0;
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        if (a * NB + b < NC * ND) {
          double v = (double)acc[a * NB + b][a * NB + b];
          mx = v > mx ? v : mx;
        }
      }
    args.schwarz[bp] = sqrt(mx);
    return;
  }

  // Cartesian -> spherical on each index (compile-time sparse coefficients).
  constexpr int SA = nsph(LA), SB = nsph(LB), SC = nsph(LC), SD = nsph(LD);
  Real s1[NA * NB][NC][SD];
#pragma unroll
  for (int i = 0; i < NA * NB; ++i)
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int sd = 0; sd < SD; ++sd) {
        if (i % NB < B0 || i % NB >= B1) continue;
        Real v = (Real)0;
#pragma unroll
        for (int d = 0; d < ND; ++d)
          if (SPH[LD][sd][d] != 0.0) v += (Real)SPH[LD][sd][d] * acc[i][c * ND + d];
        s1[i][c][sd] = v;
      }
  Real s2[NA * NB][SC * SD];
#pragma unroll
  for (int i = 0; i < NA * NB; ++i)
#pragma unroll
    for (int sc = 0; sc < SC; ++sc)
#pragma unroll
      for (int sd = 0; sd < SD; ++sd) {
        if (i % NB < B0 || i % NB >= B1) continue;
        Real v = (Real)0;
#pragma unroll
        for (int c = 0; c < NC; ++c)
          if (SPH[LC][sc][c] != 0.0) v += (Real)SPH[LC][sc][c] * s1[i][c][sd];
        s2[i][sc * SD + sd] = v;
      }
  Real s3[NA][SB][SC * SD];
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int sb = 0; sb < SB; ++sb)
#pragma unroll
      for (int k = 0; k < SC * SD; ++k) {
        if (k < CD0 || k >= CD1) continue;  // a range only with identity ket transforms
        if (sb < B0 || sb >= B1) continue;  // likewise for b
        Real v = (Real)0;
#pragma unroll
        for (int b = 0; b < NB; ++b)
          if (SPH[LB][sb][b] != 0.0) v += (Real)SPH[LB][sb][b] * s2[a * NB + b][k];
        s3[a][sb][k] = v;
      }

  const int ao_a = args.sp_ao[2 * bp], ao_b = args.sp_ao[2 * bp + 1];
  const int ao_c = args.sp_ao[2 * kp], ao_d = args.sp_ao[2 * kp + 1];
  const int64_t n = args.nbf;
#pragma unroll
  for (int sa = 0; sa < SA; ++sa)
#pragma unroll
    for (int sb = 0; sb < SB; ++sb)
#pragma unroll
      for (int k = 0; k < SC * SD; ++k) {
        if (k < CD0 || k >= CD1) continue;
        if (sb < B0 || sb >= B1) continue;
        Real v = (Real)0;
#pragma unroll
        for (int a = 0; a < NA; ++a)
          if (SPH[LA][sa][a] != 0.0) v += (Real)SPH[LA][sa][a] * s3[a][sb][k];
        const double val = (double)v;
        const int64_t I = ao_a + sa, J = ao_b + sb, K = ao_c + k / SD, Lx = ao_d + k % SD;
        if (args.mode == ERI_DENSE) {
          double* o = args.out;
          o[((I * n + J) * n + K) * n + Lx] = val;
          o[((J * n + I) * n + K) * n + Lx] = val;
          o[((I * n + J) * n + Lx) * n + K] = val;
          o[((J * n + I) * n + Lx) * n + K] = val;
          o[((K * n + Lx) * n + I) * n + J] = val;
          o[((K * n + Lx) * n + J) * n + I] = val;
          o[((Lx * n + K) * n + I) * n + J] = val;
          o[((Lx * n + K) * n + J) * n + I] = val;
        } else {
          store_super(args.out + (args.sp_base ? args.sp_base[bp] : 0), I, J, K, Lx, val);
        }
      }
}

// Launch one class; instantiated once per (class, precision) in its own
// translation unit (generated by build.py) so the classes compile in parallel.
// Ket Cartesian pairs per launch: three for FP64 (dp|pp) and one for FP64
// (dd|ps), which spill to scratch as single launches (three pairs fit (dp|pp)
// in about 440 registers, and beat nine one-pair launches at twice the
// occupancy: 3.2 against 4.5 ms on (H2O)_8); all of them otherwise (and
// always in Schwarz mode).
template <int LA, int LB, int LC, int LD, typename Real>
constexpr int eri_cd_chunk() {
  constexpr bool f64 = sizeof(Real) == 8;
  if (f64 && LA == 2 && LB == 1 && LC == 1 && LD == 1) return 3;
  if (f64 && LA == 2 && LB == 2 && LC == 1 && LD == 0) return 1;
  return ncart(LC) * ncart(LD);
}

template <int LA, int LB, int LC, int LD, typename Real, int CD0, int CH, int NCD>
void eri_launch_class_chunks(dim3 grid, int block, const EriArgs& a, hipStream_t s) {
  if constexpr (CD0 < NCD) {
    constexpr int CD1 = CD0 + CH < NCD ? CD0 + CH : NCD;
    hipLaunchKernelGGL((eri_class_kernel<LA, LB, LC, LD, Real, CD0, CD1>), grid, dim3(block), 0, s, a);
    eri_launch_class_chunks<LA, LB, LC, LD, Real, CD1, CH, NCD>(grid, block, a, s);
  }
}

// Bra Cartesians b per launch, for a p shell b: one for FP64 (dp|ds), whose
// d ket rules out ket ranges and which spills to scratch as one launch; all
// of them otherwise.
template <int LA, int LB, int LC, int LD, typename Real>
constexpr int eri_b_chunk() {
  constexpr bool f64 = sizeof(Real) == 8;
  if (f64 && LA == 2 && LB == 1 && LC == 2 && LD == 0) return 1;
  return ncart(LB);
}

template <int LA, int LB, int LC, int LD, typename Real, int B0, int NB>
void eri_launch_class_bchunks(dim3 grid, int block, const EriArgs& a, hipStream_t s) {
  if constexpr (B0 < NB) {
    hipLaunchKernelGGL((eri_class_kernel<LA, LB, LC, LD, Real, 0, ncart(LC) * ncart(LD), B0, B0 + 1>),
                       grid, dim3(block), 0, s, a);
    eri_launch_class_bchunks<LA, LB, LC, LD, Real, B0 + 1, NB>(grid, block, a, s);
  }
}

template <int LA, int LB, int LC, int LD, typename Real>
hipError_t eri_launch_class(const EriArgs& a, hipStream_t s) {
  if (a.nquartets <= 0) return hipSuccess;
  const int block = a.block == 64 ? 64 : 128;
  const dim3 grid = grid_1d((a.nquartets + block - 1) / block, block);
  constexpr int NCD = ncart(LC) * ncart(LD);
  constexpr int CH = eri_cd_chunk<LA, LB, LC, LD, Real>();
  static_assert(CH == NCD || (LC <= 1 && LD <= 1), "ket ranges need identity ket transforms");
  constexpr int BCH = eri_b_chunk<LA, LB, LC, LD, Real>();
  static_assert(BCH == ncart(LB) || (BCH == 1 && LB <= 1 && CH == NCD), "bra ranges: one b each, p shells");
  if constexpr (BCH < ncart(LB)) {
    if (a.mode != ERI_SCHWARZ) {
      eri_launch_class_bchunks<LA, LB, LC, LD, Real, 0, ncart(LB)>(grid, block, a, s);
      return hipGetLastError();
    }
  }
  if (CH < NCD && a.mode != ERI_SCHWARZ)
    eri_launch_class_chunks<LA, LB, LC, LD, Real, 0, CH, NCD>(grid, block, a, s);
  else
    hipLaunchKernelGGL((eri_class_kernel<LA, LB, LC, LD, Real, 0, NCD>), grid, dim3(block), 0, s, a);
  return hipGetLastError();
}

}  // namespace qcmd
