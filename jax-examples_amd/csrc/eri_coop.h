// eri_coop.h -- wave-cooperative McMurchie-Davidson ERIs for the classes
// where one thread per quartet would need several hundred live values (R_tuv
// alone has 165 entries at L = 8) and spills to scratch.
//
// One work-group of one or two 64-lane waves per contracted quartet; every
// thread owns a fixed set of elements of the Cartesian block V[ab][cd] and
// accumulates them in registers.  Per bra primitive pair: per bra Cartesian pair ab the list of
// its nonzero E^{ab}_{tuv} (coefficient, Hermite index).  Per batch of ket
// primitive pairs (as many as the LDS budget allows): Boys values, the
// R^n_tuv recursion (jax_qc/integrals/mcmurchie_davidson.py:113-172) level
// by level with the (primitive, tuv) entries of a level spread over the
// lanes and R^0 kept as a box so that index shifts add; per ket Cartesian
// pair cd the list of its terms pref (-1)^{t'+u'+v'} E^{cd}_{t'u'v'};
//   h_cd[tuv] = sum_{t'u'v'} c_{t'u'v'} R_{t+t',u+u',v+v'}   (t+u+v <= la+lb)
// summed over all ket primitives in registers (each thread owns fixed
// (cd, tuv) entries); then once per bra primitive pair
// V[ab][cd] += sum_{tuv} E^{ab}_{tuv} h_cd[tuv] (obara_saika.py:285-331;
// E^{ab} does not depend on the ket primitive, so the bra contraction of the
// ket-summed h equals the sum of the per-primitive ones).  After all
// primitives the block goes through LDS to the spherical basis (all lanes)
// and is written out like the per-thread kernels.
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

// (t, u, v) of Hermite index k, packed t | u << 8 | v << 16.
template <int L>
struct HermDecode {
  int tuv[nherm(L)];
  constexpr HermDecode() : tuv() {
    for (int s = 0; s <= L; ++s)
      for (int t = s; t >= 0; --t)
        for (int u = s - t; u >= 0; --u) tuv[herm_index(t, u, s - t - u)] = t | (u << 8) | ((s - t - u) << 16);
  }
};
template <int L>
__constant__ HermDecode<L> kHermDecode = HermDecode<L>();

// E^{ij}_t of a table built by hermite_e<LA, LB>.
template <int LA, int LB, typename Real>
__device__ inline Real eidx(const Real* e, int i, int j, int t) {
  return e[(i * (LB + 1) + j) * (LA + LB + 1) + t];
}

// Largest number of Hermite terms (t, u, v) of one Cartesian pair of shells
// (la, lb): prod over x, y, z of (a_x + b_x + 1).
__host__ __device__ constexpr int pair_terms(int la, int lb) {
  int m = 1;
  for (int a = 0; a < ncart(la); ++a)
    for (int b = 0; b < ncart(lb); ++b) {
      const int sx = cart_lx(la, a) + cart_lx(lb, b), sy = cart_ly(la, a) + cart_ly(lb, b);
      const int sz = la + lb - sx - sy;
      const int n = (sx + 1) * (sy + 1) * (sz + 1);
      m = n > m ? n : m;
    }
  return m;
}

// sum_j c[j] R[d[j]] over a padded term list: every offset first, then
// every R value, then the products (the loads of a list are in flight
// together instead of one load-wait-FMA step per term)
template <int N, typename Real>
__device__ __forceinline__ Real term_sum(const Real* c, const int* d, const Real* R) {
  int dv[N];
  Real rv[N];
#pragma unroll
  for (int j = 0; j < N; ++j) dv[j] = d[j];
#pragma unroll
  for (int j = 0; j < N; ++j) rv[j] = R[dv[j]];
  Real s = (Real)0;
#pragma unroll
  for (int j = 0; j < N; ++j) s += c[j] * rv[j];
  return s;
}

// A padded term list of one Cartesian pair: coefficients, then offsets.
template <int N, typename Real>
struct TermList {
  Real c[N];
  int d[N];
};

#ifndef QC_EW_KC_BUDGET
#define QC_EW_KC_BUDGET 2560
#endif
#ifndef QC_EW_KC_MIN
#define QC_EW_KC_MIN 7
#endif
#ifndef QC_EW_WAVES
#define QC_EW_WAVES 0  // > 0: that many waves per SIMD for every class (tuning sweeps)
#endif
constexpr int EW_BUDGET = 6144;                // 8-byte words of LDS per wave
constexpr int EW_KC_BUDGET = QC_EW_KC_BUDGET;  // batching ket primitives stops here (occupancy)

// One work-group (NT threads) per contracted quartet.  Per bra primitive pair: the bra
// E tables and, per bra Cartesian pair ab, the list of its nonzero
// E^{ab}_{tuv} (coefficient, Hermite index).  Per batch of ket primitive
// pairs: ket E tables, (-2 alpha)^n F_n, the R^0_tuv box (R recursion level
// by level), per ket Cartesian pair cd the list of (pref (-1)^{t'+u'+v'}
// E^{cd}_{t'u'v'}, box offset) terms, then h_cd[tuv] += sum_terms c R[tuv + t'u'v']
// (entries (cd, tuv) owned by fixed lanes, kept in registers).  After the
// last batch the ket-summed h goes to LDS and every lane adds to the
// Cartesian block elements it owns: V[ab][cd] += sum_terms E^{ab}_{tuv} h_cd[tuv].
// Both term lists are padded with zero terms to the class's largest count, so
// the term loops have compile-time trip counts and issue all their LDS loads
// back to back (a wave is latency-bound at one or two waves per SIMD).
template <int LA, int LB, int LC, int LD, typename Real>
struct EriWave {
  static constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD;
  static constexpr int NB = ncart(LB), ND = ncart(LD);
  static constexpr int NAB = ncart(LA) * NB, NCD = ncart(LC) * ND;
  static constexpr int NHB = nherm(LAB), NHR = nherm(L);
  static constexpr int EBS = (LA + 1) * (LB + 1) * (LAB + 1);
  static constexpr int EKS = (LC + 1) * (LD + 1) * (LCD + 1);
  static constexpr int RB = L + 1;
  static constexpr int TB = pair_terms(LA, LB), TK = pair_terms(LC, LD);
  // threads per quartet work-group: two waves share one quartet's LDS for
  // (dd|dd) (measured on (H2O)_8: FP64 2.12 -> 1.47 ms, and half the block
  // elements per thread, 422 -> 198 VGPRs); one wave elsewhere, where two
  // measured slower ((dp|dp) 14.6 -> 17.6 ms, (dd|dp) 7.5 -> 9.7 ms)
  static constexpr int NT = LC + LD >= 4 ? 128 : 64;
  static constexpr int IPL = (NAB * NCD + NT - 1) / NT;  // block elements per thread
  static constexpr int HPL = (NCD * NHB + NT - 1) / NT;  // h_cd[tuv] entries per thread
  static constexpr int HS = NHB | 1;  // LDS row stride of the ket-summed h (odd)
  static constexpr int WR = sizeof(Real) == 8 ? 2 : 1;  // Real per 8-byte word, inverse
  // 8-byte words: fixed, per ket primitive of a batch
  static constexpr int FIX = (3 * EBS + NAB * TB) / (3 - WR) + (NAB * (TB + 1)) / 2 + NHB / 2 + 5;
  static constexpr int KPER = (3 * EKS + 2 * NHR + RB * RB * RB + (L + 1) + 4 + NCD * TK) / (3 - WR) +
                              (NCD * (TK + 1)) / 2 + 2;
  static constexpr int BLK = (NAB * NCD > NCD * HS ? NAB * NCD : NCD * HS) / (3 - WR) + 1;
  static constexpr int lds(int kc) { return FIX + (kc * KPER > BLK ? kc * KPER : BLK); }
  static constexpr int pick_kc() {
    int kc = COOP_MAXP;
    while (kc > 1 && lds(kc) > EW_KC_BUDGET) --kc;
    return kc;
  }
  static constexpr int KC = pick_kc() >= QC_EW_KC_MIN ? pick_kc() : 1;
  // waves per SIMD the register budget must allow: 3 for (dd|dp) FP64 (225
  // VGPRs unconstrained, two waves; capped: (H2O)_16 48.4 -> 44.5 ms); every
  // other class measured slower capped at 3 ((dd|dd) FP64 7.9 -> 10.1 ms,
  // (dp|dp) 103 -> 108) and keeps the compiler's choice
  static constexpr int OCC = QC_EW_WAVES > 0 ? QC_EW_WAVES
                             : (LA == 2 && LB == 2 && LC == 2 && LD == 1 && sizeof(Real) == 8) ? 3
                                                                                                : 1;
  struct PerKet {
    // first in LDS, so that a list's coefficients and offsets sit at small
    // immediate offsets from one base address; padded to TK terms with zero
    // coefficients
    TermList<TK, Real> kt[KC][NCD];
    Real ek[KC][3][EKS];
    Real r[2][KC][NHR];
    Real rc[KC][RB * RB * RB];
    Real f[KC][L + 1];
    Real misc[KC][4];  // X, Y, Z, pref
  };
  struct Shared {
    union {
      PerKet k;
      Real hs[NCD * HS];    // ket-summed h_cd[tuv], after a bra primitive's ket loop
      Real blk[NAB * NCD];  // the contracted Cartesian block, after the primitive loops
    } u;
    Real eb[3][EBS];
    TermList<TB, Real> bt[NAB];  // padded to TB terms with zero coefficients
    Real pv[5];  // bra primitive pair: exponent, centre, prefactor
    int hbox[NHB];  // R-box offset of the (t, u, v) of Hermite index k
  };
};

template <int LA, int LB, int LC, int LD, typename Real>
__global__ __launch_bounds__((EriWave<LA, LB, LC, LD, Real>::NT))
__attribute__((amdgpu_waves_per_eu(EriWave<LA, LB, LC, LD, Real>::OCC))) void eri_coop_kernel(EriArgs args) {
  using W = EriWave<LA, LB, LC, LD, Real>;
  constexpr int LAB = W::LAB, LCD = W::LCD, L = W::L;
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int NAB = W::NAB, NCD = W::NCD, NHB = W::NHB, NHR = W::NHR, RB = W::RB;
  constexpr int TB = W::TB, TK = W::TK, KC = W::KC, IPL = W::IPL, HPL = W::HPL, HS = W::HS;
  const HermTable<L>& HT = kHermTable<L>;
  __shared__ typename W::Shared sh;
  auto& K = sh.u.k;

  constexpr int NT = W::NT;
  const int tid = threadIdx.x;
  int bp, kp;
  const int64_t gid = flat_block();
  if (gid >= args.nquartets) return;  // whole work-group (one wave)
  locate(args, gid, bp, kp);
  const Real* __restrict__ pp = pair_table<Real>(args);
  const int pb0 = args.sp_prim_off[bp], npb = args.sp_nprim[bp];
  const int pk0 = args.sp_prim_off[kp], npk = args.sp_nprim[kp];

  // Per-thread constants of the loops below, read from the constant tables
  // once: the recursion source of every R-level item the thread computes
  // (packed dir | src1 << 2 | (src2 + 1) << 10 | mult << 19), the R-box
  // offset of its level-0 items and the Cartesian exponents of its bra and
  // ket term lists (packed 2 bits each).  Inside the primitive loops these
  // lookups would each be a dependent global-memory round trip.
  constexpr int RSL = (KC * NHR + NT - 1) / NT;  // R-level items per thread
  constexpr int KSL = (KC * NCD + NT - 1) / NT;  // ket term lists per thread
  constexpr int BSL = (NAB + NT - 1) / NT;       // bra term lists per thread
  int prec[L + 1][RSL], rbx[RSL], kexp[KSL], bexp[BSL];
#pragma unroll
  for (int n = 0; n <= L; ++n)
#pragma unroll
    for (int i = 0; i < RSL; ++i) {
      prec[n][i] = 0;
      if (i * NT < KC * nherm(L - n)) {
        const int k = (tid + NT * i) % nherm(L - n);
        const HermRec hr = HT.rec[k];
        prec[n][i] = hr.dir | ((hr.src1 & 255) << 2) | ((hr.src2 + 1) << 10) | (hr.mult << 19);
      }
    }
#pragma unroll
  for (int i = 0; i < RSL; ++i) {
    const int tuv = kHermDecode<L>.tuv[(tid + NT * i) % NHR];
    rbx[i] = ((tuv & 255) * RB + ((tuv >> 8) & 255)) * RB + (tuv >> 16);
  }
#pragma unroll
  for (int i = 0; i < KSL; ++i) {
    const int cd = (tid + NT * i) % NCD, c = cd / ND, d = cd - c * ND;
    kexp[i] = CX[LC][c] | (CY[LC][c] << 2) | (CZ[LC][c] << 4) | (CX[LD][d] << 6) | (CY[LD][d] << 8) |
              (CZ[LD][d] << 10);
  }
#pragma unroll
  for (int i = 0; i < BSL; ++i) {
    const int ab = (tid + NT * i) % NAB, a = ab / NB, b = ab - a * NB;
    bexp[i] = CX[LA][a] | (CY[LA][a] << 2) | (CZ[LA][a] << 4) | (CX[LB][b] << 6) | (CY[LB][b] << 8) |
              (CZ[LB][b] << 10);
  }
  for (int k = tid; k < NHB; k += NT) {
    const int tuv = kHermDecode<LAB>.tuv[k];
    sh.hbox[k] = ((tuv & 255) * RB + ((tuv >> 8) & 255)) * RB + (tuv >> 16);
  }
  // Ket primitive-pair values a thread reads: the E lanes (tid < 3 KC: ket
  // primitive tid / 3, direction tid % 3) PA, PB and the exponent; the Boys
  // lanes (32 <= tid < 32 + KC) the exponent, the centre and the prefactor.
  // (Loading them one batch ahead measured slower on (H2O)_16: the registers
  // cost (dd|pp) FP64 a wave per SIMD, 42.6 -> 59.5 ms.)
  static_assert(3 * KC <= 32, "E lanes and Boys lanes overlap");
  const bool elane = tid < 3 * KC, blane = tid >= 32 && tid < 32 + KC;
  const int qkk = elane ? tid / 3 : tid - 32, qdim = elane ? tid - 3 * (tid / 3) : 0;
  int qf[5];
  if (elane) {
    qf[0] = PP_PAX + qdim;
    qf[1] = PP_PBX + qdim;
    qf[2] = qf[3] = qf[4] = PP_P;
  } else {
    qf[0] = PP_P;
    qf[1] = PP_PX;
    qf[2] = PP_PX + 1;
    qf[3] = PP_PX + 2;
    qf[4] = PP_K;
  }
  Real acc[IPL];
#pragma unroll
  for (int j = 0; j < IPL; ++j) acc[j] = (Real)0;

  for (int ib = 0; ib < npb; ++ib) {
    const Real* P = pp + (int64_t)(pb0 + ib) * PP_STRIDE;
    __syncthreads();
    if (tid == 3) {
      sh.pv[0] = P[PP_P];
      sh.pv[1] = P[PP_PX];
      sh.pv[2] = P[PP_PX + 1];
      sh.pv[3] = P[PP_PX + 2];
      sh.pv[4] = P[PP_K];
    }
    if (tid < 3) {
      Real E[LA + 1][LB + 1][LAB + 1];
      hermite_e<LA, LB, Real>(E, P[PP_PAX + tid], P[PP_PBX + tid], (Real)0.5 / P[PP_P]);
#pragma unroll
      for (int i = 0; i <= LA; ++i)
#pragma unroll
        for (int j = 0; j <= LB; ++j)
#pragma unroll
          for (int t = 0; t <= LAB; ++t) sh.eb[tid][(i * (LB + 1) + j) * (LAB + 1) + t] = E[i][j][t];
    }
    __syncthreads();
    // bra term lists
#pragma unroll
    for (int i = 0; i < BSL; ++i) {
      const int ab = tid + NT * i;
      if (ab < NAB) {
        int e = bexp[i];
        asm volatile("" : "+v"(e));
        const int ax = e & 3, ay = (e >> 2) & 3, az = (e >> 4) & 3;
        const int bx = (e >> 6) & 3, by = (e >> 8) & 3, bz = (e >> 10) & 3;
        int n = 0;
        for (int t = 0; t <= ax + bx; ++t) {
          const Real ex = eidx<LA, LB, Real>(sh.eb[0], ax, bx, t);
          for (int u = 0; u <= ay + by; ++u) {
            const Real exy = ex * eidx<LA, LB, Real>(sh.eb[1], ay, by, u);
            for (int v = 0; v <= az + bz; ++v) {
              sh.bt[ab].c[n] = exy * eidx<LA, LB, Real>(sh.eb[2], az, bz, v);
              sh.bt[ab].d[n] = herm_index(t, u, v);
              ++n;
            }
          }
        }
        for (; n < TB; ++n) {
          sh.bt[ab].c[n] = (Real)0;
          sh.bt[ab].d[n] = 0;
        }
      }
    }
    Real hacc[HPL];
#pragma unroll
    for (int j = 0; j < HPL; ++j) hacc[j] = (Real)0;
    for (int ik0 = 0; ik0 < npk; ik0 += KC) {
      const int nk = min(KC, npk - ik0);
      __syncthreads();
      Real qv[5];
      if ((elane || blane) && qkk < nk) {
        const Real* Q = pp + (int64_t)(pk0 + ik0 + qkk) * PP_STRIDE;
#pragma unroll
        for (int j = 0; j < 5; ++j) qv[j] = Q[qf[j]];
      }
      if (elane && qkk < nk) {
        Real E[LC + 1][LD + 1][LCD + 1];
        hermite_e<LC, LD, Real>(E, qv[0], qv[1], (Real)0.5 / qv[2]);
#pragma unroll
        for (int i = 0; i <= LC; ++i)
#pragma unroll
          for (int j = 0; j <= LD; ++j)
#pragma unroll
            for (int t = 0; t <= LCD; ++t) K.ek[qkk][qdim][(i * (LD + 1) + j) * (LCD + 1) + t] = E[i][j][t];
      }
      if (blane && qkk < nk) {
        const Real p = sh.pv[0], q = qv[0];
        const Real pqm = p * q, ps = p + q, alpha = pqm / ps;
        const Real X = sh.pv[1] - qv[1], Y = sh.pv[2] - qv[2], Z = sh.pv[3] - qv[3];
        Real F[L + 1];
        boys_eval<L, Real>(alpha * (X * X + Y * Y + Z * Z), args.boys64, F);
        Real pw = (Real)1;
#pragma unroll
        for (int n = 0; n <= L; ++n) {
          K.f[qkk][n] = pw * F[n];
          pw *= (Real)-2 * alpha;
        }
        K.misc[qkk][0] = X;
        K.misc[qkk][1] = Y;
        K.misc[qkk][2] = Z;
        K.misc[qkk][3] = (Real)34.986836655249725693 / (pqm * dev_sqrt(ps)) * sh.pv[4] * qv[4];
      }
      __syncthreads();
      // ket term lists (pref and sign folded in)
#pragma unroll
      for (int i = 0; i < KSL; ++i) {
        const int w = tid + NT * i;
        if (w < nk * NCD) {
          const int kk = w / NCD, cd = w - kk * NCD;
          int e = kexp[i];
          asm volatile("" : "+v"(e));
          const int cx = e & 3, cy = (e >> 2) & 3, cz = (e >> 4) & 3;
          const int dx = (e >> 6) & 3, dy = (e >> 8) & 3, dz = (e >> 10) & 3;
          const Real pref = K.misc[kk][3];
          int n = 0;
          for (int tt = 0; tt <= cx + dx; ++tt) {
            const Real ex = eidx<LC, LD, Real>(K.ek[kk][0], cx, dx, tt);
            for (int uu = 0; uu <= cy + dy; ++uu) {
              const Real exy = ex * eidx<LC, LD, Real>(K.ek[kk][1], cy, dy, uu);
              for (int vv = 0; vv <= cz + dz; ++vv) {
                const Real e2 = pref * exy * eidx<LC, LD, Real>(K.ek[kk][2], cz, dz, vv);
                K.kt[kk][cd].c[n] = ((tt + uu + vv) & 1) ? -e2 : e2;
                K.kt[kk][cd].d[n] = (tt * RB + uu) * RB + vv;
                ++n;
              }
            }
          }
          for (; n < TK; ++n) {
            K.kt[kk][cd].c[n] = (Real)0;
            K.kt[kk][cd].d[n] = 0;
          }
        }
      }
      // R^0_tuv (t+u+v <= L) per ket primitive of the batch, levels n = L
      // .. 0; level 0 also into the RB^3 box
#pragma unroll
      for (int n = L; n >= 0; --n) {
        const int nh = nherm(L - n);
#pragma unroll
        for (int i = 0; i < RSL; ++i) {
          const int w = tid + NT * i;
          if (i * NT < KC * nh && w < nk * nh) {
            const int kk = w / nh, k = w - kk * nh;
            Real* cur = K.r[n & 1][kk];
            const Real* prv = K.r[(n + 1) & 1][kk];
            Real v;
            if (k == 0) {
              v = K.f[kk][n];
            } else {
              int pr = prec[n][i];
              asm volatile("" : "+v"(pr));  // decode here, not hoisted out of the loops
              v = K.misc[kk][pr & 3] * prv[(pr >> 2) & 255];
              const int s2 = ((pr >> 10) & 511) - 1;
              if (s2 >= 0) v += (Real)(pr >> 19) * prv[s2];
            }
            cur[k] = v;
            if (n == 0) K.rc[kk][rbx[i]] = v;
          }
        }
        __syncthreads();
      }
      // h_cd[tuv] (t+u+v <= la+lb) of this thread's entries, summed over the batch
#pragma unroll
      for (int j = 0; j < HPL; ++j) {
        const int it = tid + NT * j;
        if (it < NCD * NHB) {
          const int cd = it / NHB, k = it - cd * NHB;
          const int hoff = sh.hbox[k];
          for (int kk = 0; kk < nk; ++kk)
            hacc[j] += term_sum<TK>(K.kt[kk][cd].c, K.kt[kk][cd].d, K.rc[kk] + hoff);
        }
      }
    }
    __syncthreads();  // every lane is done with the last batch's lists and R boxes
#pragma unroll
    for (int j = 0; j < HPL; ++j) {
      const int it = tid + NT * j;
      if (it < NCD * NHB) {
        const int cd = it / NHB, k = it - cd * NHB;
        sh.u.hs[cd * HS + k] = hacc[j];
      }
    }
    __syncthreads();
    // V[ab][cd] += sum_terms E^{ab}_{tuv} h_cd[tuv] for this thread's elements
#pragma unroll
    for (int j = 0; j < IPL; ++j) {
      const int it = tid + NT * j;
      if (it < NAB * NCD) {
        const int ab = it / NCD, cd = it - ab * NCD;
        acc[j] += term_sum<TB>(sh.bt[ab].c, sh.bt[ab].d, sh.u.hs + cd * HS);
      }
    }
  }
  __syncthreads();
  Real* blk = sh.u.blk;
#pragma unroll
  for (int j = 0; j < IPL; ++j) {
    const int it = tid + NT * j;
    if (it < NAB * NCD) blk[it] = acc[j];
  }
  __syncthreads();

  if (args.mode == ERI_SCHWARZ) {
    if (tid == 0) {
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
  for (int o = tid; o < SA * SB * SC * SD; o += NT) {
    // keeps the block's LDS reads inside the loop: hoisted, every lane would
    // hold the whole Cartesian block in registers (216 doubles for (dd|ds)),
    // which set the register count of the whole kernel
    asm volatile("" ::: "memory");
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
    const int64_t I = ao_a + sa, J = ao_b + sb, Kx = ao_c + sc, Lx = ao_d + sd;
    if (args.mode == ERI_DENSE) {
      double* out = args.out;
      out[((I * n + J) * n + Kx) * n + Lx] = val;
      out[((J * n + I) * n + Kx) * n + Lx] = val;
      out[((I * n + J) * n + Lx) * n + Kx] = val;
      out[((J * n + I) * n + Lx) * n + Kx] = val;
      out[((Kx * n + Lx) * n + I) * n + J] = val;
      out[((Kx * n + Lx) * n + J) * n + I] = val;
      out[((Lx * n + Kx) * n + I) * n + J] = val;
      out[((Lx * n + Kx) * n + J) * n + I] = val;
    } else {
      store_super(args.out + (args.sp_base ? args.sp_base[bp] : 0), I, J, Kx, Lx, val);
    }
  }
}

template <int LA, int LB, int LC, int LD, typename Real>
hipError_t eri_launch_coop(const EriArgs& a, hipStream_t s) {
  if (a.nquartets <= 0) return hipSuccess;
  if (a.nquartets > 0x7fffffffLL) return hipErrorInvalidValue;
  constexpr int NT = EriWave<LA, LB, LC, LD, Real>::NT;
  hipLaunchKernelGGL((eri_coop_kernel<LA, LB, LC, LD, Real>), grid_1d(a.nquartets, NT), dim3(NT), 0, s, a);
  return hipGetLastError();
}

}  // namespace qcmd
