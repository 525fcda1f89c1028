// eri_coop.h -- wave-cooperative McMurchie-Davidson ERIs for the high
// angular-momentum classes (total L >= 5), where one thread per quartet
// would need several hundred live FP64 values (R_tuv alone has 165 entries
// at L = 8) and spills to scratch.
//
// One 64-lane wave per contracted quartet.  Per ket primitive pair:
//   stage 1: Boys values for every bra primitive pair of the current chunk,
//            then the R^n_tuv recursion (jax_qc/integrals/mcmurchie_davidson.py:
//            113-172) run level by level with the (bra prim, tuv) entries of a
//            level spread over the lanes, levels ping-ponged through LDS;
//   stage 2: lane (cd, g) owns ket Cartesian pair cd and bra primitive pairs
//            g, g+G, ...; it forms the ket-contracted Hermite integrals
//            h_tuv = sum_{t'u'v'} (-1)^{t'+u'+v'} E^{cd}_{t'u'v'} R_{t+t',u+u',v+v'}
//            for t+u+v <= la+lb and accumulates E^{ab} . h for every bra
//            Cartesian pair ab in registers (obara_saika.py:285-331).
// After all primitives the G partial blocks of each cd are summed through
// LDS, the Cartesian block is taken to the spherical basis by all lanes, and
// written out like the per-thread kernels.
#pragma once
#include "eri_impl.h"

namespace qcmd {

constexpr int COOP_MAXP = 9;  // bra primitive pairs per LDS chunk

// Flat index of (t,u,v) in the degree-ordered tetrahedron: all entries of
// degree < s first, then (t descending, u descending).
__host__ __device__ constexpr int herm_index(int t, int u, int v) {
  const int s = t + u + v;
  return (s == 0 ? 0 : s * (s + 1) * (s + 2) / 6) + (s - t) * (s - t + 1) / 2 + (s - t - u);
}

// Recursion source table for entry k of the tetrahedron (k > 0): direction
// (0 x, 1 y, 2 z), index of the entry one step down that direction, index of
// the entry two steps down (or -1) and its multiplier.
struct HermRec {
  int dir, src1, src2, mult;
};

template <int L>
struct HermTable {
  HermRec rec[nherm(L)];
  constexpr HermTable() : rec() {
    for (int s = 0; s <= L; ++s)
      for (int t = s; t >= 0; --t)
        for (int u = s - t; u >= 0; --u) {
          const int v = s - t - u;
          const int k = herm_index(t, u, v);
          if (s == 0) {
            rec[k] = {0, -1, -1, 0};
          } else if (t > 0) {
            rec[k] = {0, herm_index(t - 1, u, v), t > 1 ? herm_index(t - 2, u, v) : -1, t - 1};
          } else if (u > 0) {
            rec[k] = {1, herm_index(t, u - 1, v), u > 1 ? herm_index(t, u - 2, v) : -1, u - 1};
          } else {
            rec[k] = {2, herm_index(t, u, v - 1), v > 1 ? herm_index(t, u, v - 2) : -1, v - 1};
          }
        }
  }
};

template <int L>
__constant__ HermTable<L> kHermTable = HermTable<L>();

template <int LA, int LB, int LC, int LD, typename Real>
struct CoopShared {
  static constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD;
  static constexpr int ESA = (LA + 1) * (LB + 1) * (LAB + 1);
  static constexpr int ESC = (LC + 1) * (LD + 1) * (LCD + 1);
  static constexpr int NHL = nherm(L);
  static constexpr int NAB = ncart(LA) * ncart(LB), NCD = ncart(LC) * ncart(LD);
  static constexpr int RED = 64 * NAB > NAB * NCD ? 64 * NAB : NAB * NCD;
  static constexpr int RB = 2 * COOP_MAXP * NHL > RED ? 2 * COOP_MAXP * NHL : RED;
  Real eb[COOP_MAXP][3][ESA];
  Real ek[3][ESC];
  Real f[COOP_MAXP][L + 1];  // (-2 alpha)^n F_n
  Real pq[COOP_MAXP][4];     // X, Y, Z, prefactor
  Real r[RB];                // two R levels, later the reduction buffer
};

template <int LA, int LB, typename Real>
__device__ inline Real eidx(const Real* e, int i, int j, int t) {
  return e[(i * (LB + 1) + j) * (LA + LB + 1) + t];
}

template <int LA, int LB, int LC, int LD, typename Real>
__global__ __launch_bounds__(64) void eri_coop_kernel(EriArgs args) {
  using S = CoopShared<LA, LB, LC, LD, Real>;
  constexpr int LAB = S::LAB, LCD = S::LCD, L = S::L;
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int NAB = S::NAB, NCD = S::NCD, NHL = S::NHL, ESA = S::ESA, ESC = S::ESC;
  constexpr int G = 64 / NCD;  // lanes per ket Cartesian pair
  const HermTable<L>& HT = kHermTable<L>;
  __shared__ S sh;

  const int lane = threadIdx.x;
  int bp, kp;
  const int64_t gid = flat_block();
  if (gid >= args.nquartets) return;  // whole work-group (one wave)
  locate(args, gid, bp, kp);
  const Real* __restrict__ pp = pair_table<Real>(args);
  const Real* __restrict__ tab = boys_table<Real>(args);
  const int pb0 = args.sp_prim_off[bp], npb = args.sp_nprim[bp];
  const int pk0 = args.sp_prim_off[kp], npk = args.sp_nprim[kp];

  const int cd = lane % NCD, g = lane / NCD;
  const bool active = g < G;
  const int c = cd / ND, d = cd % ND;
  const int cx = CX[LC][c] , cy = CY[LC][c], cz = CZ[LC][c];
  const int dx = CX[LD][d], dy = CY[LD][d], dz = CZ[LD][d];
  const int sx = cx + dx, sy = cy + dy, sz = cz + dz;

  Real acc[NAB];
#pragma unroll
  for (int i = 0; i < NAB; ++i) acc[i] = (Real)0;

  for (int ib0 = 0; ib0 < npb; ib0 += COOP_MAXP) {
    const int nb = min(COOP_MAXP, npb - ib0);
    __syncthreads();
    if (lane < nb) {
      const Real* P = pp + (int64_t)(pb0 + ib0 + lane) * PP_STRIDE;
      const Real oo2p = (Real)0.5 / P[PP_P];
      Real E[LA + 1][LB + 1][LAB + 1];
#pragma unroll
      for (int dim = 0; dim < 3; ++dim) {
        hermite_e<LA, LB, Real>(E, P[PP_PAX + dim], P[PP_PBX + dim], oo2p);
#pragma unroll
        for (int i = 0; i <= LA; ++i)
#pragma unroll
          for (int j = 0; j <= LB; ++j)
#pragma unroll
            for (int t = 0; t <= LAB; ++t) sh.eb[lane][dim][(i * (LB + 1) + j) * (LAB + 1) + t] = E[i][j][t];
      }
    }
    for (int ik = 0; ik < npk; ++ik) {
      const Real* Q = pp + (int64_t)(pk0 + ik) * PP_STRIDE;
      const Real q = Q[PP_P];
      __syncthreads();
      if (lane < 3) {
        Real E[LC + 1][LD + 1][LCD + 1];
        hermite_e<LC, LD, Real>(E, Q[PP_PAX + lane], Q[PP_PBX + lane], (Real)0.5 / q);
#pragma unroll
        for (int i = 0; i <= LC; ++i)
#pragma unroll
          for (int j = 0; j <= LD; ++j)
#pragma unroll
            for (int t = 0; t <= LCD; ++t) sh.ek[lane][(i * (LD + 1) + j) * (LCD + 1) + t] = E[i][j][t];
      }
      if (lane >= 32 && lane - 32 < nb) {
        const int ib = lane - 32;
        const Real* P = pp + (int64_t)(pb0 + ib0 + ib) * PP_STRIDE;
        const Real p = P[PP_P];
        const Real pqm = p * q, ps = p + q, alpha = pqm / ps;
        const Real X = P[PP_PX] - Q[PP_PX], Y = P[PP_PX + 1] - Q[PP_PX + 1],
                   Z = P[PP_PX + 2] - Q[PP_PX + 2];
        Real F[L + 1];
        boys_fn<L, Real>(alpha * (X * X + Y * Y + Z * Z), tab, F);
        Real pw = (Real)1;
#pragma unroll
        for (int n = 0; n <= L; ++n) {
          sh.f[ib][n] = pw * F[n];
          pw *= (Real)-2 * alpha;
        }
        sh.pq[ib][0] = X;
        sh.pq[ib][1] = Y;
        sh.pq[ib][2] = Z;
        sh.pq[ib][3] = (Real)34.986836655249725693 / (pqm * dev_sqrt(ps)) * P[PP_K] * Q[PP_K];
      }
      __syncthreads();
      // stage 1: R^n for n = L .. 0, level n in buffer n & 1
#pragma unroll
      for (int n = L; n >= 0; --n) {
        const int nh = nherm(L - n);
        Real* cur = sh.r + (n & 1) * COOP_MAXP * NHL;
        const Real* prv = sh.r + ((n + 1) & 1) * COOP_MAXP * NHL;
        for (int e = lane; e < nb * nh; e += 64) {
          const int ib = e / nh, k = e - ib * nh;
          Real v;
          if (k == 0) {
            v = sh.f[ib][n];
          } else {
            const HermRec hr = HT.rec[k];
            v = sh.pq[ib][hr.dir] * prv[ib * NHL + hr.src1];
            if (hr.src2 >= 0) v += (Real)hr.mult * prv[ib * NHL + hr.src2];
          }
          cur[ib * NHL + k] = v;
        }
        __syncthreads();
      }
      // stage 2
      if (active) {
        const Real* R0 = sh.r;  // level 0
        for (int ib = g; ib < nb; ib += G) {
          const Real* Rb = R0 + ib * NHL;
          Real h[nherm(LAB)];
#pragma unroll
          for (int t = 0; t <= LAB; ++t)
#pragma unroll
            for (int u = 0; u <= LAB - t; ++u)
#pragma unroll
              for (int v = 0; v <= LAB - t - u; ++v) {
                Real s = (Real)0;
                for (int tt = 0; tt <= sx; ++tt) {
                  const Real ex = eidx<LC, LD, Real>(sh.ek[0], cx, dx, tt);
                  for (int uu = 0; uu <= sy; ++uu) {
                    const Real exy = ex * eidx<LC, LD, Real>(sh.ek[1], cy, dy, uu);
                    for (int vv = 0; vv <= sz; ++vv) {
                      const Real e = exy * eidx<LC, LD, Real>(sh.ek[2], cz, dz, vv);
                      const Real r = Rb[herm_index(t + tt, u + uu, v + vv)];
                      s += ((tt + uu + vv) & 1) ? -e * r : e * r;
                    }
                  }
                }
                h[herm_index(t, u, v)] = s;
              }
          const Real pref = sh.pq[ib][3];
          const Real* ebx = sh.eb[ib][0];
          const Real* eby = sh.eb[ib][1];
          const Real* ebz = sh.eb[ib][2];
#pragma unroll
          for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int b = 0; b < NB; ++b) {
              const int ax = CX[LA][a], ay = CY[LA][a], az = CZ[LA][a];
              const int bx = CX[LB][b], by = CY[LB][b], bz = CZ[LB][b];
              Real val = (Real)0;
#pragma unroll
              for (int t = 0; t <= LAB; ++t) {
                if (t > ax + bx) continue;
#pragma unroll
                for (int u = 0; u <= LAB; ++u) {
                  if (u > ay + by) continue;
                  const Real exy = eidx<LA, LB, Real>(ebx, ax, bx, t) * eidx<LA, LB, Real>(eby, ay, by, u);
#pragma unroll
                  for (int v = 0; v <= LAB; ++v) {
                    if (v > az + bz) continue;
                    val += exy * eidx<LA, LB, Real>(ebz, az, bz, v) * h[herm_index(t, u, v)];
                  }
                }
              }
              acc[a * NB + b] += pref * val;
            }
        }
      }
    }
  }
  (void)ESA;
  (void)ESC;

  // sum the G partial blocks of each cd: red[lane][ab]
  __syncthreads();
  Real* red = sh.r;
  if (active) {
#pragma unroll
    for (int i = 0; i < NAB; ++i) red[lane * NAB + i] = acc[i];
  }
  __syncthreads();
  // Cartesian block blk[ab][cd] (NAB x NCD), built by lanes cd < NCD
  Real tot[NAB];
  if (lane < NCD) {
#pragma unroll
    for (int i = 0; i < NAB; ++i) tot[i] = red[lane * NAB + i];
    for (int gg = 1; gg < G; ++gg)
#pragma unroll
      for (int i = 0; i < NAB; ++i) tot[i] += red[(gg * NCD + lane) * NAB + i];
  }
  __syncthreads();
  Real* blk = sh.r;
  if (lane < NCD) {
#pragma unroll
    for (int i = 0; i < NAB; ++i) blk[i * NCD + lane] = tot[i];
  }
  __syncthreads();

  if (args.mode == ERI_SCHWARZ) {
    if (lane == 0) {
      double mx = 0.0;
      for (int i = 0; i < NAB && i < NCD; ++i) {
        const double v = (double)blk[i * NCD + i];
        mx = v > mx ? v : mx;
      }
      args.schwarz[bp] = sqrt(mx);
    }
    return;
  }

  constexpr int SA = nsph(LA), SB = nsph(LB), SC = nsph(LC), SD = nsph(LD);
  const int ao_a = args.sp_ao[2 * bp], ao_b = args.sp_ao[2 * bp + 1];
  const int ao_c = args.sp_ao[2 * kp], ao_d = args.sp_ao[2 * kp + 1];
  const int64_t n = args.nbf;
  for (int o = lane; o < SA * SB * SC * SD; o += 64) {
    const int sa = o / (SB * SC * SD), sb = (o / (SC * SD)) % SB, sc = (o / SD) % SC, sd = o % SD;
    Real v = (Real)0;
    for (int a = 0; a < NA; ++a) {
      const Real ca = (Real)SPH[LA][sa][a];
      if (ca == (Real)0) continue;
      for (int b = 0; b < NB; ++b) {
        const Real cab = ca * (Real)SPH[LB][sb][b];
        if (cab == (Real)0) continue;
        for (int cc = 0; cc < NC; ++cc) {
          const Real cabc = cab * (Real)SPH[LC][sc][cc];
          if (cabc == (Real)0) continue;
          for (int dd = 0; dd < ND; ++dd) {
            const Real w = cabc * (Real)SPH[LD][sd][dd];
            if (w != (Real)0) v += w * blk[(a * NB + b) * NCD + cc * ND + dd];
          }
        }
      }
    }
    const double val = (double)v;
    const int64_t I = ao_a + sa, J = ao_b + sb, K = ao_c + sc, Lx = ao_d + sd;
    if (args.mode == ERI_DENSE) {
      double* out = args.out;
      out[((I * n + J) * n + K) * n + Lx] = val;
      out[((J * n + I) * n + K) * n + Lx] = val;
      out[((I * n + J) * n + Lx) * n + K] = val;
      out[((J * n + I) * n + Lx) * n + K] = val;
      out[((K * n + Lx) * n + I) * n + J] = val;
      out[((K * n + Lx) * n + J) * n + I] = val;
      out[((Lx * n + K) * n + I) * n + J] = val;
      out[((Lx * n + K) * n + J) * n + I] = val;
    } else {
      store_super(args.out + (args.sp_base ? args.sp_base[bp] : 0), I, J, K, Lx, val);
    }
  }
}

template <int LA, int LB, int LC, int LD, typename Real>
hipError_t eri_launch_coop(const EriArgs& a, hipStream_t s) {
  if (a.nquartets <= 0) return hipSuccess;
  if (a.nquartets > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((eri_coop_kernel<LA, LB, LC, LD, Real>), grid_1d(a.nquartets, 64),
                     dim3(64), 0, s, a);
  return hipGetLastError();
}

}  // namespace qcmd
