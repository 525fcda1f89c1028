// grad_impl.h -- two-electron part of the analytic nuclear gradient,
// contracted on the fly (no derivative integrals are stored).
//
//   dE_2e/dX = 1/2 sum_{canonical shell quartets} f sum_{abcd} G_abcd d(ab|cd)/dX
//   G_abcd   = Dj_ab Dj_cd - cx sum_s (Dk_s,ac Dk_s,bd + Dk_s,ad Dk_s,bc)
// (RHF: Dj = Dk = D, cx = 1/4; UHF: Dj = Da + Db, Dk = {Da, Db}, cx = 1/2),
// f = the number of distinct images of the shell quartet, all matrices in
// the Cartesian AO basis.  Derivatives follow from the Gaussian identity
// d/dA_x x_A^a e^{-alpha x_A^2} = 2 alpha x_A^{a+1} e^{..} - a x_A^{a-1} e^{..}
// applied to the 1-D Hermite coefficients (E^A_t = 2 alpha E^{a+1,b}_t -
// a E^{a-1,b}_t), for centres A, B and C; D by translational invariance.
// This is the contraction the reference's rhf_gradient forms by finite
// differences of full integral rebuilds (jax_qc/grad/rhf_grad.py:153-209).
//
// Two kernel forms (chosen per class): one thread per contracted quartet,
// everything in registers (ket (s s) classes), and one 64-lane wave per
// quartet with the bra side contracted with G first (the rest); both add the
// quartet's A, B, C, D contributions atomically into grad[atom][xyz].
#pragma once
#include "eri_coop.h"

namespace qcmd {

struct Grad2eArgs {
  const double* pairs64;
  const int* sp_prim_off;
  const int* sp_nprim;
  const int* sp_aoc;   // per shell pair: Cartesian AO offsets (shell 1, shell 2)
  const int* sp_atom;  // per shell pair: atoms (shell 1, shell 2)
  const int* sp_diag;  // per shell pair: 1 when both shells are the same
  const int* bra_list;
  const int* ket_list;
  const int64_t* offsets;
  const int* kstart;
  int nbra;
  int64_t nquartets;
  int64_t ncart;
  const double* dj;
  const double* dk0;
  const double* dk1;
  int nk;
  double cx;
  double* grad;
  int natom;  // atoms in grad (0: not given)
  const double* boys64;
  const double* schwarz;  // per shell pair Q (density screening), or null
  double tau_density;
  const float* pairs32;  // FP32 launches: primitive-pair table and Boys grid
  const float* boys32;
};

template <typename Real>
__device__ inline const Real* gpairs(const Grad2eArgs& a);
template <>
__device__ inline const double* gpairs<double>(const Grad2eArgs& a) { return a.pairs64; }
template <>
__device__ inline const float* gpairs<float>(const Grad2eArgs& a) { return a.pairs32; }
template <typename Real>
__device__ inline const Real* gboys(const Grad2eArgs& a);
template <>
__device__ inline const double* gboys<double>(const Grad2eArgs& a) { return a.boys64; }
template <>
__device__ inline const float* gboys<float>(const Grad2eArgs& a) { return a.boys32; }

constexpr int GMAXP = 9;  // ket primitive pairs per batch, at most

// One thread per contracted quartet, for the low classes (the G block and
// the 9 accumulators stay in registers).  Per primitive quartet: E tables up
// to (la+1, lb+1) and their centre derivatives 2 alpha E^{a+1} - a E^{a-1},
// R up to L+1; for each ket Cartesian pair cd the ket-contracted Hermite
// integrals h0_tuv (t+u+v <= la+lb+1) and, per direction, hC_tuv with the
// ket's C-derivative coefficients (t+u+v <= la+lb); each bra pair ab then
// takes its A-, B- and undifferentiated E against them.
// Ket Cartesian pairs cd in [CD0, CD1) only (one launch per range adds its
// share of the gradient): a range keeps the live registers of the FP64
// (pp|ps) and (ds|ps) kernels within the register file.
template <int LA, int LB, int LC, int LD, typename Real, int CD0, int CD1>
__device__ __forceinline__ bool grad2e_thread_quartet(const Grad2eArgs& args, int64_t gid, double (&g)[12],
                                                      int (&at)[4]) {
  constexpr int LAB = LA + LB, LCD = LC + LD, L1 = LAB + LCD + 1;
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int NAB = NA * NB, NCD = NC * ND;
  constexpr int NH1 = nherm(LAB + 1), NH0 = nherm(LAB);
  int lo = 0, hi = args.nbra - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (args.offsets[mid] <= gid)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int bp = args.bra_list[lo];
  const int kp = args.ket_list[args.kstart[lo] + (int)(gid - args.offsets[lo])];
  const Real* __restrict__ pp = gpairs<Real>(args);
  const int pb0 = args.sp_prim_off[bp], npb = args.sp_nprim[bp];
  const int pk0 = args.sp_prim_off[kp], npk = args.sp_nprim[kp];
  const int64_t oa = args.sp_aoc[2 * bp], ob = args.sp_aoc[2 * bp + 1];
  const int64_t oc = args.sp_aoc[2 * kp], od = args.sp_aoc[2 * kp + 1];
  const int64_t nc = args.ncart;

  // G over the launch's cd range; with density screening, max |G| over the
  // whole block (every range's launch takes the same decision)
  const bool screen = args.tau_density > 0.0;
  Real gam[NAB][NCD];
  double gmax = 0.0;
#pragma unroll
  for (int ab = 0; ab < NAB; ++ab)
#pragma unroll
    for (int cd = 0; cd < NCD; ++cd) {
      const bool mine = cd >= CD0 && cd < CD1;
      if (!mine && !screen) continue;
      const int64_t mu = oa + ab / NB, nu = ob + ab % NB, la = oc + cd / ND, si = od + cd % ND;
      double k = args.dk0[mu * nc + la] * args.dk0[nu * nc + si] +
                 args.dk0[mu * nc + si] * args.dk0[nu * nc + la];
      if (args.nk > 1)
        k += args.dk1[mu * nc + la] * args.dk1[nu * nc + si] +
             args.dk1[mu * nc + si] * args.dk1[nu * nc + la];
      const Real gv = (Real)(args.dj[mu * nc + nu] * args.dj[la * nc + si] - args.cx * k);
      if (mine) gam[ab][cd] = gv;
      gmax = fmax(gmax, fabs((double)gv));
    }
  if (screen && gmax * args.schwarz[bp] * args.schwarz[kp] < args.tau_density) return false;

  Real acc[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) acc[i] = (Real)0;

  for (int ik = 0; ik < npk; ++ik) {
    const Real* Q = pp + (int64_t)(pk0 + ik) * PP_STRIDE;
    const Real q = Q[PP_P], ga2 = (Real)2 * Q[PP_ALPHA];
    // ket tables up to (lc+1, ld): the C derivative only (D by invariance)
    Real Ek[3][LC + 2][LD + 1][LCD + 2];
    Real dEk[3][LC + 1][LD + 1][LCD + 2];
#pragma unroll
    for (int dim = 0; dim < 3; ++dim) {
      hermite_e<LC + 1, LD, Real>(Ek[dim], Q[PP_PAX + dim], Q[PP_PBX + dim], (Real)0.5 / q);
#pragma unroll
      for (int i = 0; i <= LC; ++i)
#pragma unroll
        for (int j = 0; j <= LD; ++j)
#pragma unroll
          for (int t = 0; t <= LCD + 1; ++t) {
            Real v = ga2 * Ek[dim][i + 1][j][t];
            if (i > 0) v -= (Real)i * Ek[dim][i > 0 ? i - 1 : 0][j][t];
            dEk[dim][i][j][t] = v;
          }
    }
    for (int ib = 0; ib < npb; ++ib) {
      const Real* P = pp + (int64_t)(pb0 + ib) * PP_STRIDE;
      const Real p = P[PP_P], al2 = (Real)2 * P[PP_ALPHA];
      // bra tables up to (la+1, lb) and the A derivative only: B follows from
      // the Hermite shift d/dP R_tuv = R_{t+1,u,v} (dB = dP - dA)
      Real Eb[3][LA + 2][LB + 1][LAB + 2];
      Real dEa[3][LA + 1][LB + 1][LAB + 2];
#pragma unroll
      for (int dim = 0; dim < 3; ++dim) {
        hermite_e<LA + 1, LB, Real>(Eb[dim], P[PP_PAX + dim], P[PP_PBX + dim], (Real)0.5 / p);
#pragma unroll
        for (int i = 0; i <= LA; ++i)
#pragma unroll
          for (int j = 0; j <= LB; ++j)
#pragma unroll
            for (int t = 0; t <= LAB + 1; ++t) {
              Real va = al2 * Eb[dim][i + 1][j][t];
              if (i > 0) va -= (Real)i * Eb[dim][i > 0 ? i - 1 : 0][j][t];
              dEa[dim][i][j][t] = va;
            }
      }
      const Real pqm = p * q, ps = p + q, alpha = pqm / ps;
      const Real X = P[PP_PX] - Q[PP_PX], Y = P[PP_PX + 1] - Q[PP_PX + 1],
                   Z = P[PP_PX + 2] - Q[PP_PX + 2];
      Real F[L1 + 1];
      boys_eval<L1, Real>(alpha * (X * X + Y * Y + Z * Z), args.boys64, F);
      Real R[L1 + 1][L1 + 1][L1 + 1];
      hermite_r<L1, Real>(R, F, (Real)-2 * alpha, X, Y, Z);
      const Real pref = (Real)34.986836655249725693 / (pqm * dev_sqrt(ps)) * P[PP_K] * Q[PP_K];

#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int d = 0; d < ND; ++d) {
          if (c * ND + d < CD0 || c * ND + d >= CD1) continue;
          const int kx = CX[LC][c] + 0, ky = CY[LC][c], kz = CZ[LC][c];
          const int jx = CX[LD][d], jy = CY[LD][d], jz = CZ[LD][d];
          Real h0[NH1], hC[3][NH0];
#pragma unroll
          for (int t = 0; t <= LAB + 1; ++t)
#pragma unroll
            for (int u = 0; u <= LAB + 1 - t; ++u)
#pragma unroll
              for (int v = 0; v <= LAB + 1 - t - u; ++v) {
                const bool low = t + u + v <= LAB;
                Real s0 = (Real)0, sx = (Real)0, sy = (Real)0, sz = (Real)0;
#pragma unroll
                for (int tt = 0; tt <= kx + jx + 1; ++tt)
#pragma unroll
                  for (int uu = 0; uu <= ky + jy + 1; ++uu)
#pragma unroll
                    for (int vv = 0; vv <= kz + jz + 1; ++vv) {
                      const int over = (tt > kx + jx) + (uu > ky + jy) + (vv > kz + jz);
                      if (over > 1) continue;
                      Real r = R[t + tt][u + uu][v + vv];
                      if ((tt + uu + vv) & 1) r = -r;
                      const Real ex0 = Ek[0][kx][jx][tt], ey0 = Ek[1][ky][jy][uu],
                                   ez0 = Ek[2][kz][jz][vv];
                      if (over == 0) s0 += ex0 * ey0 * ez0 * r;
                      if (low) {
                        if (uu <= ky + jy && vv <= kz + jz) sx += dEk[0][kx][jx][tt] * ey0 * ez0 * r;
                        if (tt <= kx + jx && vv <= kz + jz) sy += ex0 * dEk[1][ky][jy][uu] * ez0 * r;
                        if (tt <= kx + jx && uu <= ky + jy) sz += ex0 * ey0 * dEk[2][kz][jz][vv] * r;
                      }
                    }
                h0[herm_index(t, u, v)] = s0;
                if (low) {
                  hC[0][herm_index(t, u, v)] = sx;
                  hC[1][herm_index(t, u, v)] = sy;
                  hC[2][herm_index(t, u, v)] = sz;
                }
              }
#pragma unroll
          for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int b = 0; b < NB; ++b) {
              const int ax = CX[LA][a], ay = CY[LA][a], az = CZ[LA][a];
              const int bx = CX[LB][b], by = CY[LB][b], bz = CZ[LB][b];
              Real s[9];
#pragma unroll
              for (int i = 0; i < 9; ++i) s[i] = (Real)0;
#pragma unroll
              for (int t = 0; t <= ax + bx + 1; ++t)
#pragma unroll
                for (int u = 0; u <= ay + by + 1; ++u)
#pragma unroll
                  for (int v = 0; v <= az + bz + 1; ++v) {
                    const int over = (t > ax + bx) + (u > ay + by) + (v > az + bz);
                    if (over > 1) continue;
                    const Real e0x = Eb[0][ax][bx][t], e0y = Eb[1][ay][by][u], e0z = Eb[2][az][bz][v];
                    const Real hv = h0[herm_index(t, u, v)];
                    if (u <= ay + by && v <= az + bz) s[0] += dEa[0][ax][bx][t] * e0y * e0z * hv;
                    if (t <= ax + bx && v <= az + bz) s[1] += e0x * dEa[1][ay][by][u] * e0z * hv;
                    if (t <= ax + bx && u <= ay + by) s[2] += e0x * e0y * dEa[2][az][bz][v] * hv;
                    if (over == 0) {
                      const Real e0 = e0x * e0y * e0z;
                      // P = A + B: the bra distribution moved along x, y, z
                      s[3] += e0 * h0[herm_index(t + 1, u, v)];
                      s[4] += e0 * h0[herm_index(t, u + 1, v)];
                      s[5] += e0 * h0[herm_index(t, u, v + 1)];
                      s[6] += e0 * hC[0][herm_index(t, u, v)];
                      s[7] += e0 * hC[1][herm_index(t, u, v)];
                      s[8] += e0 * hC[2][herm_index(t, u, v)];
                    }
                  }
              const Real gm = gam[a * NB + b][c * ND + d] * pref;
#pragma unroll
              for (int i = 0; i < 9; ++i) acc[i] += gm * s[i];
            }
        }
    }
  }
  const double f = (args.sp_diag[bp] ? 1.0 : 2.0) * (args.sp_diag[kp] ? 1.0 : 2.0) *
                   (bp == kp ? 1.0 : 2.0) * 0.5;
  at[0] = args.sp_atom[2 * bp];
  at[1] = args.sp_atom[2 * bp + 1];
  at[2] = args.sp_atom[2 * kp];
  at[3] = args.sp_atom[2 * kp + 1];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double gA = f * acc[k], gP = f * acc[3 + k], gC = f * acc[6 + k];
    g[k] = gA;
    g[3 + k] = gP - gA;
    g[6 + k] = gC;
    g[9 + k] = -(gP + gC);
  }
  return true;
}

// atoms whose gradient a thread-per-quartet work-group sums in LDS before
// adding it to the global gradient (one global add per touched component
// and work-group instead of twelve per quartet)
constexpr int GT_LDS_ATOMS = 256;

template <int LA, int LB, int LC, int LD, typename Real, int CD0, int CD1>
__global__ __launch_bounds__(128) void grad2e_thread_kernel(Grad2eArgs args) {
  __shared__ double gsh[3 * GT_LDS_ATOMS];
  const bool staged = args.natom > 0 && args.natom <= GT_LDS_ATOMS;  // uniform
  const int n3 = 3 * args.natom;
  if (staged) {
    for (int i = threadIdx.x; i < n3; i += blockDim.x) gsh[i] = 0.0;
    __syncthreads();
  }
  const int64_t gid = flat_block() * blockDim.x + threadIdx.x;
  double g[12];
  int at[4];
  const bool on = gid < args.nquartets && grad2e_thread_quartet<LA, LB, LC, LD, Real, CD0, CD1>(args, gid, g, at);
  if (staged) {
    if (on) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          // an atom index at or past natom (a caller's wrong count) goes
          // straight to the global gradient instead of outside gsh
          if (at[i] < args.natom)
            atomicAdd(&gsh[3 * at[i] + k], g[3 * i + k]);
          else
            atomicAdd(&args.grad[3 * at[i] + k], g[3 * i + k]);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n3; i += blockDim.x)
      if (gsh[i] != 0.0) atomicAdd(&args.grad[i], gsh[i]);
  } else if (on) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) atomicAdd(&args.grad[3 * at[i] + k], g[3 * i + k]);
  }
}

// Ket Cartesian pairs per launch of the thread-per-quartet kernel: one where
// all of them in one launch would not fit the registers well (FP64 (pp|ps)
// and (ds|ps) need about 800 and spill: 1.01 -> 0.17 s and 0.17 -> 0.10 s on
// (H2O)_16 per pair; (ds|ds) FP64 fits only per pair, 0.03 s against 0.35 s
// for the wave form; FP32 (pp|ps) 52 -> 45 ms), all of them otherwise.
template <int LA, int LB, int LC, int LD, typename Real>
constexpr int thread_cd_chunk() {
  constexpr bool f64 = sizeof(Real) == 8;
  if (LC == 1 && LD == 0 && LA + LB == 2) return 1;
  if (f64 && LA == 2 && LB == 0 && LC == 2 && LD == 0) return 1;
  return ncart(LC) * ncart(LD);
}

template <int LA, int LB, int LC, int LD, typename Real, int CD0, int CH, int NCD>
void grad2e_launch_thread_chunks(dim3 grid, const Grad2eArgs& a, hipStream_t s) {
  if constexpr (CD0 < NCD) {
    constexpr int CD1 = CD0 + CH < NCD ? CD0 + CH : NCD;
    hipLaunchKernelGGL((grad2e_thread_kernel<LA, LB, LC, LD, Real, CD0, CD1>), grid, dim3(128), 0, s, a);
    grad2e_launch_thread_chunks<LA, LB, LC, LD, Real, CD1, CH, NCD>(grid, a, s);
  }
}

template <int LA, int LB, int LC, int LD, typename Real = double>
hipError_t grad2e_launch_thread(const Grad2eArgs& a, hipStream_t s) {
  if (a.nquartets <= 0) return hipSuccess;
  const dim3 grid = grid_1d((a.nquartets + 127) / 128, 128);
  constexpr int NCD = ncart(LC) * ncart(LD);
  grad2e_launch_thread_chunks<LA, LB, LC, LD, Real, 0, thread_cd_chunk<LA, LB, LC, LD, Real>(), NCD>(grid, a, s);
  return hipGetLastError();
}



// ---------------------------------------------------------------------------
// One wave per contracted quartet with the bra side contracted with G first.
// Per bra primitive pair (alpha, p) and ket Cartesian pair cd:
//   Y0_cd[tuv]    = sum_ab G_abcd E^{ab}_{tuv}                  (t+u+v <= la+lb)
//   YA_cd,x[tuv]  = sum_ab G_abcd dE^{ab}/dA_x _{tuv}           (t+u+v <= la+lb+1)
// with dE^{ab}/dA_x = 2 alpha E^{a+1,b} - a_x E^{a-1,b}.  Per primitive
// quartet, with h_cd[tuv] = sum_{t'u'v'} (-1)^{t'+u'+v'} E^{cd}_{t'u'v'} R_{t+t',u+u',v+v'}
// and hC_cd,x the same with the ket's C-derivative coefficients:
//   dA_x  = pref sum_cd sum_tuv YA_cd,x[tuv] h_cd[tuv]
//   dP_x  = pref sum_cd sum_tuv Y0_cd[tuv] h_cd[t+1,u,v]     (d/dA + d/dB: the
//           bra distribution moves with P and d/dP_x R_{tuv} = R_{t+1,u,v})
//   dC_x  = pref sum_cd sum_tuv Y0_cd[tuv] hC_cd,x[tuv]
// then dB = dP - dA and dD = -(dP + dC) (translational invariance).  Y is
// built once per bra primitive and reused by every ket primitive; the work
// items (tuv, cd) are spread over all lanes of the work-group (one or two
// waves, NT) and ket pairs are taken in
// chunks sized so that the LDS stays under GW_BUDGET doubles.
// ---------------------------------------------------------------------------

#ifndef QC_GW_BUDGET
#define QC_GW_BUDGET 6144
#endif
constexpr int GW_BUDGET = QC_GW_BUDGET;  // doubles of LDS per wave (48 KB)
#ifndef QC_GW_KC_BUDGET
#define QC_GW_KC_BUDGET 2560
#endif
#ifndef QC_GW_KC_MIN
#define QC_GW_KC_MIN 7
#endif
constexpr int GW_KC_BUDGET = QC_GW_KC_BUDGET;  // batching ket primitives stops at 20 KB (8 waves per CU)

// Largest number of Hermite terms of one ket Cartesian pair (with one extra
// order in some direction for the C-derivative lists).
__host__ __device__ constexpr int tmax_terms(int lc, int ld) {
  int m = 1;
  for (int c = 0; c < ncart(lc); ++c)
    for (int d = 0; d < ncart(ld); ++d) {
      const int sx = cart_lx(lc, c) + cart_lx(ld, d), sy = cart_ly(lc, c) + cart_ly(ld, d);
      const int sz = lc + ld - sx - sy;
      const int n0 = (sx + 2) * (sy + 1) * (sz + 1), n1 = (sx + 1) * (sy + 2) * (sz + 1),
                n2 = (sx + 1) * (sy + 1) * (sz + 2);
      m = n0 > m ? n0 : m;
      m = n1 > m ? n1 : m;
      m = n2 > m ? n2 : m;
    }
  return m;
}

template <int LA, int LB, int LC, int LD, typename Real>
struct GradWave {
  static constexpr int LAB = LA + LB, LCD = LC + LD, L1 = LAB + LCD + 1;
  static constexpr int NB = ncart(LB), ND = ncart(LD);
  static constexpr int NAB = ncart(LA) * NB, NCD = ncart(LC) * ND;
  static constexpr int NHB0 = nherm(LAB), NHB1 = nherm(LAB + 1), NHR = nherm(L1);
  static constexpr int EBS = (LA + 2) * (LB + 1) * (LAB + 2);  // E^{ij}_t: i <= la+1, j <= lb
  static constexpr int EKS = (LC + 2) * (LD + 1) * (LCD + 2);
  static constexpr int RB = L1 + 1;  // R^0 kept as an RB^3 box: index shifts are additive
  // most ket terms (t', u', v') of one Cartesian pair cd: list 0 for h, list
  // 1 + x for hC_x (one more term along x)
  static constexpr int TMAX = tmax_terms(LC, LD);
  static constexpr int T0 = pair_terms(LC, LD);  // list 0 has no extra order
  // threads per quartet work-group: two waves share one quartet's LDS for
  // (dp) and (dd) kets, whose LDS allows few work-groups per CU (measured on
  // (H2O)_16: (dd|dp) 361 -> 294 ms, (dd|dd) 86 -> 65 ms, (dp|dp) FP32 153 ->
  // 128 ms); one wave for lighter kets, where two lost ((ds|pp) 1.19 -> 1.81 s)
  static constexpr int NT = LC + LD >= 3 ? 128 : 64;
  // LDS bytes: fixed part, per ket pair of a chunk, per ket primitive of a
  // batch, per (ket pair, ket primitive)
  static constexpr int RS = sizeof(Real);
  static constexpr int FIX = (NAB * NCD + 3 * EBS) * RS + NHB1 * 4 + 64;
  static constexpr int YPER = 4 * NHB1 * RS;
  static constexpr int KPER = (2 * NHR + RB * RB * RB) * RS;  // R levels and box
  static constexpr int EPER = (3 * EKS + (L1 + 1) + 4) * RS;   // ket E tables, F_n, X Y Z pref
  static constexpr int TPER = 4 * TMAX * (RS + 4);
  static constexpr int lds(int nc, int kc, int ke) {
    return FIX + nc * YPER + kc * (KPER + nc * TPER) + ke * EPER;
  }
  // as many ket primitives per batch (at most GMAXP) as keep the whole ket
  // side within GW_KC_BUDGET (occupancy); ket pairs are then chunked to fit
  // GW_BUDGET.  A batch shares the barrier-separated phases (ket E, Boys,
  // term lists, R levels) among its primitive quartets.
  static constexpr int pick_kc() {
    int kc = GMAXP;
    while (kc > 1 && lds(NCD, kc, kc) > 8 * GW_KC_BUDGET) --kc;
    return kc;
  }
  // batches of 2-6 cost more in occupancy than they save in barriers
  // (measured): below 7 the kernel takes one ket primitive at a time
  static constexpr int KC = pick_kc() >= QC_GW_KC_MIN ? pick_kc() : 1;
  static constexpr int pick_nc() {
    int nc = NCD;
    while (nc > 1 && lds(nc, KC, KC) > 8 * GW_BUDGET) --nc;
    const int nchunk = (NCD + nc - 1) / nc;
    // a single chunk above 32 KB is split in two when the halves fit 32 KB:
    // (dp|dp) goes from 3 to 5 work-groups per CU and from 1.25 to 0.77 s on
    // (H2O)_16, while a third chunk or more costs more in repeated ket-side
    // work than it gains ((dd|dp), (dd|dd) measured slower at 32 KB)
    if (nchunk == 1 && lds(NCD, KC, KC) > 32768 && lds((NCD + 1) / 2, KC, KC) <= 32768) return (NCD + 1) / 2;
    return (NCD + nchunk - 1) / nchunk;  // chunks of equal size
  }
  static constexpr int NCDC = pick_nc();
  // The ket-side scalars (E tables, (-2 alpha)^n F_n, X Y Z pref: a few
  // lanes' serial work per ket primitive) are built for KE ket primitives
  // at once.  With KC = 1, KE is the largest count (at most GMAXP) that keeps
  // the work-groups per CU (160 KB of LDS in 512-byte granules; the estimate
  // adds 1 KB for alignment padding) at min(that of KE = 1, 8); for (pp)
  // kets at that of KE = 1 exactly (measured: giving up occupancy for
  // batching lost in every (pp)-ket class, keeping it gains up to 5 %).
  static constexpr int occ(int b) { return 163840 / ((b + 1024 + 511) / 512 * 512); }
  static constexpr int pick_ke() {
    if (KC > 1) return KC;
    const int occ1 = occ(lds(NCDC, 1, 1));
    const int want = LC == 1 && LD == 1 ? occ1 : (occ1 < 8 ? occ1 : 8);
    int ke = GMAXP;
    while (ke > 1 && occ(lds(NCDC, 1, ke)) < want) --ke;
    return ke;
  }
  static constexpr int KE = pick_ke();
  // one ket primitive's term lists for one ket pair cd: coefficients (sign,
  // pref and derivative factors folded in) and offsets in the R box; list 0
  // padded with zero terms to T0, lists 1-3 to TMAX, so that the term loops
  // have compile-time trip counts
  struct Terms {
    Real c[4][TMAX];
    int d[4][TMAX];
  };
  struct Shared {
    // first in LDS: an item's coefficients and offsets then sit at small
    // immediate offsets from one base address
    Terms tl[KC][NCDC];
    Real gam[NAB * NCD];
    Real eb[3][EBS];
    // per Hermite index k < NHB1 of (t, u, v): the R-box offset of (t, u, v)
    // in bits 0-9 and the Hermite indices of (t-1, u, v), (t, u-1, v),
    // (t, u, v-1) in bits 10-16, 17-23, 24-30 (NHB1 - 1 where the lowered
    // index does not exist: y0 is zero there, t + u + v = la + lb + 1)
    int hk[NHB1];
    Real y0[NCDC][NHB1];
    Real ya[3][NCDC][NHB1];
    Real ek[KE][3][EKS];
    Real r[2][KC][NHR];
    Real rc[KC][RB * RB * RB];
    Real f[KE][L1 + 1];
    Real misc[KE][4];           // X, Y, Z, pref
    double wmax[NT / 64];  // per-wave max |G| (density screening)
    // MFMA form of the Y phase: per bra Cartesian pair ab, the Cartesian
    // exponents a_x, a_y, a_z in bits 10-11, 12-13, 14-15 and b_x, b_y, b_z
    // in bits 16-17, 18-19, 20-21
    int abinfo[NAB];
  };
};

// The Y phase as matrix products on the matrix cores (FP64: v_mfma_f64_
// 16x16x4f64; FP32: v_mfma_f32_16x16x4f32):
//   Y_o[k][cl] = sum_ab A_o[k][ab] G[ab][cd0 + cl],  o = 0 (E^{ab}_k),
//   1..3 (dE^{ab}_k / dA_x,y,z)
// with the Hermite index k as the row (16 per tile), the chunk's ket pair cl
// as the column (16 per tile) and ab as the reduction (4 per step).  Lane l
// supplies A[16 kt + l % 16][4 s + l / 16] (all four o from one set of E
// reads) and B[4 s + l / 16][16 ct + l % 16], and holds the four results
// D[16 kt + row(l / 16, r)][16 ct + l % 16], r < 4, with row(g, r) = 4 g + r
// for the FP32 instruction and g + 4 r for the FP64 one (both read off the
// device with one-hot inputs).
template <typename Real>
struct MfmaAcc;
template <>
struct MfmaAcc<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 step(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int r) { return g + 4 * r; }
};
template <>
struct MfmaAcc<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 step(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};

// Waves per SIMD that the LDS use of a class allows (4 SIMDs and 160 KB of
// LDS per CU), at most 3 (at most 2 for (dd| bras, whose Y phase spills
// within 3 waves' 168 registers): the register budget the compiler schedules
// for, so that registers do not lower the occupancy below what LDS allows.
template <int LA, int LB, int LC, int LD, typename Real>
constexpr int grad_wave_occupancy() {
  using W = GradWave<LA, LB, LC, LD, Real>;
  const int w = (int)(163840 / ((sizeof(typename W::Shared) + 511) / 512 * 512)) * (W::NT / 64) / 4;
  const int cap = LA + LB == 4 ? 2 : 3;
  return w < 1 ? 1 : (w > cap ? cap : w);
}

template <int LA, int LB, int LC, int LD, typename Real, bool YM = false>
__global__ __launch_bounds__((GradWave<LA, LB, LC, LD, Real>::NT)) __attribute__((amdgpu_waves_per_eu(grad_wave_occupancy<LA, LB, LC, LD, Real>())))
void grad2e_wave_kernel(Grad2eArgs args) {
  using W = GradWave<LA, LB, LC, LD, Real>;
  constexpr int LAB = W::LAB, LCD = W::LCD, L1 = W::L1, NB = W::NB, ND = W::ND;
  constexpr int NAB = W::NAB, NCD = W::NCD, NHB0 = W::NHB0, NHB1 = W::NHB1, NHR = W::NHR;
  constexpr int NCDC = W::NCDC, KC = W::KC, KE = W::KE, RB = W::RB, TMAX = W::TMAX, T0 = W::T0;
  constexpr int NT = W::NT;  // threads of the work-group
  __shared__ typename W::Shared sh;
  const HermTable<L1>& HT = kHermTable<L1>;
  const HermDecode<LAB + 1>& HD = kHermDecode<LAB + 1>;

  const int tid = threadIdx.x, lane = tid & 63;
  const int64_t gid = flat_block();
  if (gid >= args.nquartets) return;  // whole work-group (one wave)
  int lo = 0, hi = args.nbra - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (args.offsets[mid] <= gid)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int bp = args.bra_list[lo];
  const int kp = args.ket_list[args.kstart[lo] + (int)(gid - args.offsets[lo])];
  const Real* __restrict__ pp = gpairs<Real>(args);
  const int pb0 = args.sp_prim_off[bp], npb = args.sp_nprim[bp];
  const int pk0 = args.sp_prim_off[kp], npk = args.sp_nprim[kp];
  const int64_t oa = args.sp_aoc[2 * bp], ob = args.sp_aoc[2 * bp + 1];
  const int64_t oc = args.sp_aoc[2 * kp], od = args.sp_aoc[2 * kp + 1];
  const int64_t nc = args.ncart;

  for (int e = tid; e < NAB * NCD; e += NT) {
    const int ab = e / NCD, cd = e % NCD;
    const int64_t mu = oa + ab / NB, nu = ob + ab % NB, la = oc + cd / ND, si = od + cd % ND;
    double k = args.dk0[mu * nc + la] * args.dk0[nu * nc + si] +
               args.dk0[mu * nc + si] * args.dk0[nu * nc + la];
    if (args.nk > 1)
      k += args.dk1[mu * nc + la] * args.dk1[nu * nc + si] +
           args.dk1[mu * nc + si] * args.dk1[nu * nc + la];
    sh.gam[e] = (Real)(args.dj[mu * nc + nu] * args.dj[la * nc + si] - args.cx * k);
  }
  if (args.tau_density > 0.0) {
    // each thread's maximum over the G entries it wrote, then over the wave
    // and the work-group
    double gmax = 0.0;
    for (int e = tid; e < NAB * NCD; e += NT) gmax = fmax(gmax, fabs((double)sh.gam[e]));
    for (int w = 32; w > 0; w >>= 1) gmax = fmax(gmax, __shfl_xor(gmax, w));
    if (NT > 64) {
      if (lane == 0) sh.wmax[tid >> 6] = gmax;
      __syncthreads();
      gmax = fmax(sh.wmax[0], sh.wmax[NT / 64 - 1]);
    }
    if (gmax * args.schwarz[bp] * args.schwarz[kp] < args.tau_density) return;  // uniform
  }

  // E^{ij}_t of the bra / ket tables in LDS (zero for t > i + j).
  auto eb = [&](int dim, int i, int j, int t) -> Real {
    return i < 0 ? (Real)0 : sh.eb[dim][(i * (LB + 1) + j) * (LAB + 2) + t];
  };
  auto ek = [&](int kk, int dim, int i, int j, int t) -> Real {
    return i < 0 ? (Real)0 : sh.ek[kk][dim][(i * (LD + 1) + j) * (LCD + 2) + t];
  };

  for (int k = tid; k < NHB1; k += NT) {
    const int tuv = HD.tuv[k], t = tuv & 255, u = (tuv >> 8) & 255, v = tuv >> 16;
    const int none = NHB1 - 1;
    const int it = t > 0 ? herm_index(t - 1, u, v) : none, iu = u > 0 ? herm_index(t, u - 1, v) : none,
              iv = v > 0 ? herm_index(t, u, v - 1) : none;
    sh.hk[k] = ((t * RB + u) * RB + v) | (it << 10) | (iu << 17) | (iv << 24);
  }
  if constexpr (YM) {
    for (int ab = tid; ab < NAB; ab += NT) {
      const int a = ab / NB, b = ab - a * NB;
      const int ax = CX[LA][a], ay = CY[LA][a], az = CZ[LA][a];
      const int bx = CX[LB][b], by = CY[LB][b], bz = CZ[LB][b];
      // the three directions' tables share one layout: (i (LB+1) + j) (LAB+2)
      sh.abinfo[ab] = (ax << 10) | (ay << 12) | (az << 14) | (bx << 16) | (by << 18) | (bz << 20);
    }
  }
  // n / d for 0 <= n < 2^11, 1 <= d <= 2^9 by one multiply: inv = ceil(2^20 / d)
  // (exact over that range; n = ket primitive x ket pair of a chunk, d = ket pairs)
  static_assert(KC * NCDC < 2048 && NCDC <= 512, "div_small range");
  auto div_small = [](int n, unsigned inv) -> int { return (int)(((unsigned)n * inv) >> 20); };

  Real acc[9];  // dA xyz, dP xyz, dC xyz
#pragma unroll
  for (int i = 0; i < 9; ++i) acc[i] = (Real)0;

  for (int ib = 0; ib < npb; ++ib) {
    const Real* P = pp + (int64_t)(pb0 + ib) * PP_STRIDE;
    __syncthreads();
    if (tid < 3) {
      Real E[LA + 2][LB + 1][LAB + 2];
      hermite_e<LA + 1, LB, Real>(E, P[PP_PAX + tid], P[PP_PBX + tid], (Real)0.5 / P[PP_P]);
#pragma unroll
      for (int i = 0; i <= LA + 1; ++i)
#pragma unroll
        for (int j = 0; j <= LB; ++j)
#pragma unroll
          for (int t = 0; t <= LAB + 1; ++t) sh.eb[tid][(i * (LB + 1) + j) * (LAB + 2) + t] = E[i][j][t];
    }
    const Real al2 = (Real)2 * P[PP_ALPHA];
    for (int cd0 = 0; cd0 < NCD; cd0 += NCDC) {
      const int ncdc = min(NCDC, NCD - cd0);
      __syncthreads();
      if constexpr (YM) {
        using MA = MfmaAcc<Real>;
        typedef typename MA::v4 v4;
        constexpr int KT = (NHB1 + 15) / 16, ES = (LAB + 2), JS = (LB + 1) * (LAB + 2);
        const int ct_n = (ncdc + 15) / 16;
        const int q = lane >> 4, c16 = lane & 15;
        for (int tile = tid >> 6; tile < KT * ct_n; tile += NT / 64) {
          const int kt = tile % KT, ct = tile / KT;
          const int krow = kt * 16 + c16;  // this lane's A row
          const int ccol = ct * 16 + c16;  // this lane's B column
          const bool kin = krow < NHB1, cin = ccol < ncdc;
          int t = 0, u = 0, v = 0;
          if (kin) {
            const int off = sh.hk[krow] & 1023;
            t = off / (RB * RB);
            u = (off / RB) % RB;
            v = off % RB;
          }
          v4 y0 = {0, 0, 0, 0}, yx = y0, yy = y0, yz = y0;
#pragma unroll
          for (int ab0 = 0; ab0 < NAB; ab0 += 4) {
            const int ab = ab0 + q;
            Real a0 = (Real)0, ax_ = (Real)0, ay_ = (Real)0, az_ = (Real)0, g = (Real)0;
            if (ab < NAB) {
              const int inf = sh.abinfo[ab];
              const int ax = (inf >> 10) & 3, ay = (inf >> 12) & 3, az = (inf >> 14) & 3;
              const int bx = (inf >> 16) & 3, by = (inf >> 18) & 3, bz = (inf >> 20) & 3;
              if (cin) g = sh.gam[ab * NCD + cd0 + ccol];
              if (kin) {
                const Real* ex = sh.eb[0] + bx * ES + t;
                const Real* ey = sh.eb[1] + by * ES + u;
                const Real* ez = sh.eb[2] + bz * ES + v;
                const Real e0x = ex[ax * JS], e0y = ey[ay * JS], e0z = ez[az * JS];
                Real dx = al2 * ex[(ax + 1) * JS], dy = al2 * ey[(ay + 1) * JS], dz = al2 * ez[(az + 1) * JS];
                if (ax > 0) dx -= (Real)ax * ex[(ax - 1) * JS];
                if (ay > 0) dy -= (Real)ay * ey[(ay - 1) * JS];
                if (az > 0) dz -= (Real)az * ez[(az - 1) * JS];
                const Real eyz = e0y * e0z;
                a0 = e0x * eyz;
                ax_ = dx * eyz;
                ay_ = e0x * dy * e0z;
                az_ = e0x * e0y * dz;
              }
            }
            y0 = MA::step(a0, g, y0);
            yx = MA::step(ax_, g, yx);
            yy = MA::step(ay_, g, yy);
            yz = MA::step(az_, g, yz);
          }
          if (cin) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int k = kt * 16 + MA::row(q, r);
              if (k < NHB1) {
                sh.y0[ccol][k] = y0[r];
                sh.ya[0][ccol][k] = yx[r];
                sh.ya[1][ccol][k] = yy[r];
                sh.ya[2][ccol][k] = yz[r];
              }
            }
          }
        }
      } else
      // Y for this bra primitive and ket chunk: per item (cd, tuv), the
      // E columns at (t, u, v) in registers, then every ab unrolled
      for (int it = tid; it < ncdc * NHB1; it += NT) {
        const int cl = it / NHB1, k = it - cl * NHB1, cd = cd0 + cl;
        const int off = sh.hk[k] & 1023;
        const int t = off / (RB * RB), u = (off / RB) % RB, v = off % RB;
        Real ex[LA + 2][LB + 1], ey[LA + 2][LB + 1], ez[LA + 2][LB + 1];
#pragma unroll
        for (int i = 0; i <= LA + 1; ++i)
#pragma unroll
          for (int j = 0; j <= LB; ++j) {
            ex[i][j] = sh.eb[0][(i * (LB + 1) + j) * (LAB + 2) + t];
            ey[i][j] = sh.eb[1][(i * (LB + 1) + j) * (LAB + 2) + u];
            ez[i][j] = sh.eb[2][(i * (LB + 1) + j) * (LAB + 2) + v];
          }
        Real y0 = (Real)0, yx = (Real)0, yy = (Real)0, yz = (Real)0;
#pragma unroll
        for (int ab = 0; ab < NAB; ++ab) {
          const int a = ab / NB, b = ab - a * NB;
          const int ax = CX[LA][a], ay = CY[LA][a], az = CZ[LA][a];
          const int bx = CX[LB][b], by = CY[LB][b], bz = CZ[LB][b];
          const Real g = sh.gam[ab * NCD + cd];
          const Real e0x = ex[ax][bx], e0y = ey[ay][by], e0z = ez[az][bz];
          // d/dA of the a factor: 2 alpha E^{a+1} - a E^{a-1}
          Real eax = al2 * ex[ax + 1][bx], eay = al2 * ey[ay + 1][by], eaz = al2 * ez[az + 1][bz];
          if (ax > 0) eax -= (Real)ax * ex[ax > 0 ? ax - 1 : 0][bx];
          if (ay > 0) eay -= (Real)ay * ey[ay > 0 ? ay - 1 : 0][by];
          if (az > 0) eaz -= (Real)az * ez[az > 0 ? az - 1 : 0][bz];
          const Real gx = g * e0x, gyz = g * (e0y * e0z);
          y0 += gyz * e0x;
          yx += gyz * eax;
          yy += gx * e0z * eay;
          yz += gx * e0y * eaz;
        }
        sh.y0[cl][k] = y0;  // zero for t + u + v = la + lb + 1
        sh.ya[0][cl][k] = yx;
        sh.ya[1][cl][k] = yy;
        sh.ya[2][cl][k] = yz;
      }
      const unsigned inv_nc = ((1u << 20) + ncdc - 1) / ncdc;
      for (int ie0 = 0; ie0 < npk; ie0 += KE) {
        const int ne = min(KE, npk - ie0);
        __syncthreads();
        // ket E tables (lanes 3 e + dim) and Boys values (lanes 32 + e) of
        // KE ket primitives
        if (tid < 3 * ne) {
          const int e = tid / 3, dim = tid - 3 * e;
          const Real* Q = pp + (int64_t)(pk0 + ie0 + e) * PP_STRIDE;
          Real E[LC + 2][LD + 1][LCD + 2];
          hermite_e<LC + 1, LD, Real>(E, Q[PP_PAX + dim], Q[PP_PBX + dim], (Real)0.5 / Q[PP_P]);
#pragma unroll
          for (int i = 0; i <= LC + 1; ++i)
#pragma unroll
            for (int j = 0; j <= LD; ++j)
#pragma unroll
              for (int t = 0; t <= LCD + 1; ++t) sh.ek[e][dim][(i * (LD + 1) + j) * (LCD + 2) + t] = E[i][j][t];
        }
        if (tid >= 32 && tid - 32 < ne) {
          const int e = tid - 32;
          const Real* Q = pp + (int64_t)(pk0 + ie0 + e) * PP_STRIDE;
          const Real p = P[PP_P], q = Q[PP_P];
          const Real pqm = p * q, ps = p + q, alpha = pqm / ps;
          const Real X = P[PP_PX] - Q[PP_PX], Y = P[PP_PX + 1] - Q[PP_PX + 1],
                     Z = P[PP_PX + 2] - Q[PP_PX + 2];
          Real F[L1 + 1];
          boys_eval<L1, Real>(alpha * (X * X + Y * Y + Z * Z), args.boys64, F);
          Real pw = (Real)1;
#pragma unroll
          for (int n = 0; n <= L1; ++n) {
            sh.f[e][n] = pw * F[n];
            pw *= (Real)-2 * alpha;
          }
          sh.misc[e][0] = X;
          sh.misc[e][1] = Y;
          sh.misc[e][2] = Z;
          sh.misc[e][3] = (Real)34.986836655249725693 / (pqm * dev_sqrt(ps)) * P[PP_K] * Q[PP_K];
        }
        for (int ik0 = ie0; ik0 < ie0 + ne; ik0 += KC) {
          const int nk = min(KC, ie0 + ne - ik0), kb = ik0 - ie0;
          __syncthreads();
          // ket term lists per (ket primitive, list, ket pair): list 0 for h,
          // 1 + x for hC_x; pref, the sign (-1)^{t'+u'+v'} and the derivative
          // factors are folded into the coefficients
          for (int w = tid; w < nk * 4 * ncdc; w += NT) {
            const int kk = w / (4 * ncdc), r0 = w - kk * 4 * ncdc, l = r0 / ncdc, cl = r0 - l * ncdc;
            const int cd = cd0 + cl;
            const int c = cd / ND, d = cd - c * ND;
            const int cx = CX[LC][c], cy = CY[LC][c], cz = CZ[LC][c];
            const int dx = CX[LD][d], dy = CY[LD][d], dz = CZ[LD][d];
            const Real pref = sh.misc[kb + kk][3];
            const Real ga2 = (Real)2 * pp[(int64_t)(pk0 + ik0 + kk) * PP_STRIDE + PP_ALPHA];
            const int dim = l - 1;
            const int ex2 = cx + dx + (dim == 0), ey2 = cy + dy + (dim == 1), ez2 = cz + dz + (dim == 2);
            int n = 0;
            for (int tt = 0; tt <= ex2; ++tt) {
              const Real e1 = dim == 0 ? ga2 * ek(kb + kk, 0, cx + 1, dx, tt) - (Real)cx * ek(kb + kk, 0, cx - 1, dx, tt)
                                       : ek(kb + kk, 0, cx, dx, tt);
              for (int uu = 0; uu <= ey2; ++uu) {
                const Real e2 = e1 * (dim == 1 ? ga2 * ek(kb + kk, 1, cy + 1, dy, uu) - (Real)cy * ek(kb + kk, 1, cy - 1, dy, uu)
                                               : ek(kb + kk, 1, cy, dy, uu));
                for (int vv = 0; vv <= ez2; ++vv) {
                  const Real e3 = e2 * (dim == 2 ? ga2 * ek(kb + kk, 2, cz + 1, dz, vv) - (Real)cz * ek(kb + kk, 2, cz - 1, dz, vv)
                                                 : ek(kb + kk, 2, cz, dz, vv));
                  sh.tl[kk][cl].c[l][n] = ((tt + uu + vv) & 1) ? -pref * e3 : pref * e3;
                  sh.tl[kk][cl].d[l][n] = (tt * RB + uu) * RB + vv;
                  ++n;
                }
              }
            }
            for (const int nmax = l == 0 ? T0 : TMAX; n < nmax; ++n) {
              sh.tl[kk][cl].c[l][n] = (Real)0;
              sh.tl[kk][cl].d[l][n] = 0;
            }
          }
          // R^0_tuv (t+u+v <= L+1) of every primitive quartet of the batch,
          // levels n = L+1 .. 0, ping-pong buffers; level 0 also goes into the
          // RB^3 boxes
#pragma unroll
          for (int n = L1; n >= 0; --n) {
            const int nh = nherm(L1 - n);
            for (int w = tid; w < nk * nh; w += NT) {
              const int kk = w / nh, k = w - kk * nh;
              Real* cur = sh.r[n & 1][kk];
              const Real* prv = sh.r[(n + 1) & 1][kk];
              Real v;
              if (k == 0) {
                v = sh.f[kb + kk][n];
              } else {
                const HermRec hr = HT.rec[k];
                v = sh.misc[kb + kk][hr.dir] * prv[hr.src1];
                if (hr.src2 >= 0) v += (Real)hr.mult * prv[hr.src2];
              }
              cur[k] = v;
              if (n == 0) {
                const int tuv = kHermDecode<L1>.tuv[k];
                sh.rc[kk][((tuv & 255) * RB + ((tuv >> 8) & 255)) * RB + (tuv >> 16)] = v;
              }
            }
            __syncthreads();
          }
          // list 0 (h) for every (ket primitive, cd, tuv), t + u + v <= la + lb + 1
          for (int it = tid; it < nk * ncdc * NHB1; it += NT) {
            const int q = it / NHB1, k = it - q * NHB1;
            const int kk = div_small(q, inv_nc), cl = q - kk * ncdc;
            const int hk = sh.hk[k];
            const Real* Rb = sh.rc[kk] + (hk & 1023);
            const Real h = term_sum<T0>(sh.tl[kk][cl].c[0], sh.tl[kk][cl].d[0], Rb);
            acc[0] += sh.ya[0][cl][k] * h;
            acc[1] += sh.ya[1][cl][k] * h;
            acc[2] += sh.ya[2][cl][k] * h;
            acc[3] += sh.y0[cl][(hk >> 10) & 127] * h;
            acc[4] += sh.y0[cl][(hk >> 17) & 127] * h;
            acc[5] += sh.y0[cl][hk >> 24] * h;
          }
          // lists 1-3 (hC_x, hC_y, hC_z) for t + u + v <= la + lb
          for (int it = tid; it < nk * ncdc * NHB0; it += NT) {
            const int q = it / NHB0, k = it - q * NHB0;
            const int kk = div_small(q, inv_nc), cl = q - kk * ncdc;
            const Real* Rb = sh.rc[kk] + (sh.hk[k] & 1023);
            const Real y0 = sh.y0[cl][k];
#pragma unroll
            for (int dim = 0; dim < 3; ++dim) {
              acc[6 + dim] += y0 * term_sum<TMAX>(sh.tl[kk][cl].c[1 + dim], sh.tl[kk][cl].d[1 + dim], Rb);
            }
          }
        }
      }
    }
  }
  double accd[9];  // each wave's sum in FP64 whatever the arithmetic precision
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    accd[i] = (double)acc[i];
    for (int w = 32; w > 0; w >>= 1) accd[i] += __shfl_xor(accd[i], w);
  }
  if (lane == 0) {
    const double f = (args.sp_diag[bp] ? 1.0 : 2.0) * (args.sp_diag[kp] ? 1.0 : 2.0) *
                     (bp == kp ? 1.0 : 2.0) * 0.5;
    const int A = args.sp_atom[2 * bp], B = args.sp_atom[2 * bp + 1];
    const int C = args.sp_atom[2 * kp], D = args.sp_atom[2 * kp + 1];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double gA = f * accd[k], gP = f * accd[3 + k], gC = f * accd[6 + k];
      atomicAdd(&args.grad[3 * A + k], gA);
      atomicAdd(&args.grad[3 * B + k], gP - gA);
      atomicAdd(&args.grad[3 * C + k], gC);
      atomicAdd(&args.grad[3 * D + k], -(gP + gC));
    }
  }
}

template <int LA, int LB, int LC, int LD, typename Real = double, bool YM = false>
hipError_t grad2e_launch_wave(const Grad2eArgs& a, hipStream_t s) {
  if (a.nquartets <= 0) return hipSuccess;
  constexpr int NT = GradWave<LA, LB, LC, LD, Real>::NT;
  hipLaunchKernelGGL((grad2e_wave_kernel<LA, LB, LC, LD, Real, YM>), grid_1d(a.nquartets, NT), dim3(NT), 0, s, a);
  return hipGetLastError();
}

}  // namespace qcmd
