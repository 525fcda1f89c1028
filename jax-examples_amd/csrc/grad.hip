// grad.hip -- one-electron part of the analytic nuclear gradient and the
// C-ABI entry points of the gradient kernels.
//
// E_1e = sum_{mu nu} D_mu,nu (T + V)_mu,nu - sum W_mu,nu S_mu,nu over all AO
// pairs, i.e. (2 - delta_ij) times the shell-pair block sum for i >= j.  Per
// primitive pair and Cartesian component (jax_qc/integrals/obara_saika.py:
// 70-222 for the undifferentiated forms):
//   d/dA_x X_ab = 2 alpha X_{a+1x,b} - a_x X_{a-1x,b},  X in {S, T, V^C}
//   d/dB_x X_ab = 2 beta  X_{a,b+1x} - b_x X_{a,b-1x}
//   two-centre S, T: d/dB = -d/dA;  V^C: d/dC = -(d/dA + d/dB)
// (the last is the Hellmann-Feynman term of nucleus C).  D and W are the
// Cartesian-basis densities.  Accumulated atomically into grad[atom][xyz].
#include "grad_impl.h"
#include "qcmd_internal.h"

namespace qcmd {

template <int LA, int LB>
__global__ __launch_bounds__(128) void grad1e_kernel(
    const double* __restrict__ pp, const int* __restrict__ sp_prim_off,
    const int* __restrict__ sp_nprim, const int* __restrict__ sp_aoc,
    const int* __restrict__ sp_atom, const int* __restrict__ sp_diag,
    const int* __restrict__ list, int npairs, int natom, const double* __restrict__ axyz,
    const double* __restrict__ az, int64_t nc, const double* __restrict__ Dm,
    const double* __restrict__ Wm, double* __restrict__ grad, const double* __restrict__ boys) {
  constexpr int NA = ncart(LA), NB = ncart(LB), L = LA + LB;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= npairs) return;
  const int sp = list[idx];
  const int p0 = sp_prim_off[sp], np = sp_nprim[sp];
  const int64_t oa = sp_aoc[2 * sp], ob = sp_aoc[2 * sp + 1];
  const int atA = sp_atom[2 * sp], atB = sp_atom[2 * sp + 1];
  const double fac = sp_diag[sp] ? 1.0 : 2.0;
  double dmat[NA][NB], wmat[NA][NB];
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      dmat[a][b] = fac * Dm[(oa + a) * nc + ob + b];
      wmat[a][b] = fac * Wm[(oa + a) * nc + ob + b];
    }
  double gA[3] = {0, 0, 0};  // S and T parts (B = -A)
  double vA[3] = {0, 0, 0}, vB[3] = {0, 0, 0};
  // the nuclei are split over blockIdx.y (one per thread up to 64 atoms); the
  // S and T terms belong to y == 0.  A nucleus's term is summed over the
  // primitive pairs in registers and added once (the threads of a block all
  // add to the same three components: one atomic per primitive pair made
  // them contend)
  const bool st = blockIdx.y == 0;

  for (int c0 = blockIdx.y, first = 1; first || c0 < natom; c0 += gridDim.y, first = 0) {
  double vC[3] = {0.0, 0.0, 0.0};
  for (int ip = blockIdx.z; ip < np; ip += gridDim.z) {  // primitive pairs over z (few-pair classes)
    const double* P = pp + (int64_t)(p0 + ip) * PP_STRIDE;
    const double p = P[PP_P], K = P[PP_K], al = P[PP_ALPHA], be = p - al;
    const double oo2p = 0.5 / p;
    // 1-D tables up to (LA+2, LB+1): overlaps and kinetic for a+-1
    double Ex[LA + 3][LB + 2][L + 4], Ey[LA + 3][LB + 2][L + 4], Ez[LA + 3][LB + 2][L + 4];
    hermite_e<LA + 2, LB + 1, double>(Ex, P[PP_PAX], P[PP_PBX], oo2p);
    hermite_e<LA + 2, LB + 1, double>(Ey, P[PP_PAX + 1], P[PP_PBX + 1], oo2p);
    hermite_e<LA + 2, LB + 1, double>(Ez, P[PP_PAX + 2], P[PP_PBX + 2], oo2p);
    const double fS = sqrt(3.14159265358979323846 / p);
    double S1[3][LA + 3][LB + 2];
#pragma unroll
    for (int i = 0; i <= LA + 2; ++i)
#pragma unroll
      for (int j = 0; j <= LB + 1; ++j) {
        S1[0][i][j] = fS * Ex[i][j][0];
        S1[1][i][j] = fS * Ey[i][j][0];
        S1[2][i][j] = fS * Ez[i][j][0];
      }
    // kinetic 1-D for i <= LA+1, j <= LB (obara_saika.py:115-138)
    double T1[3][LA + 2][LB + 1];
#pragma unroll
    for (int dd = 0; dd < 3; ++dd)
#pragma unroll
      for (int i = 0; i <= LA + 1; ++i)
#pragma unroll
        for (int j = 0; j <= LB; ++j) {
          double t = 4.0 * al * be * S1[dd][i + 1][j + 1];
          if (i > 0 && j > 0) t += i * j * S1[dd][i - 1][j - 1];
          if (j > 0) t -= 2.0 * al * j * S1[dd][i + 1][j - 1];
          if (i > 0) t -= 2.0 * be * i * S1[dd][i - 1][j + 1];
          T1[dd][i][j] = 0.5 * t;
        }
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        if (!st || !first) continue;
        const int ai[3] = {CX[LA][a], CY[LA][a], CZ[LA][a]};
        const int bi[3] = {CX[LB][b], CY[LB][b], CZ[LB][b]};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          // derivative direction k on the bra: S and T of (a +- 1_k, b)
          double sp_ = 1.0, sm = 1.0, tp = 0.0, tm = 0.0;
          // products over the three dimensions with dimension k shifted
          double Sp[3], Sm[3], Tp[3], Tm[3];
#pragma unroll
          for (int dd = 0; dd < 3; ++dd) {
            const int i0 = ai[dd] + (dd == k ? 1 : 0);
            const int im = ai[dd] - (dd == k ? 1 : 0);
            Sp[dd] = S1[dd][i0][bi[dd]];
            Tp[dd] = T1[dd][i0][bi[dd]];
            Sm[dd] = im >= 0 ? S1[dd][im][bi[dd]] : 0.0;
            Tm[dd] = im >= 0 ? T1[dd][im][bi[dd]] : 0.0;
          }
          sp_ = Sp[0] * Sp[1] * Sp[2];
          sm = Sm[0] * Sm[1] * Sm[2];
          tp = Tp[0] * Sp[1] * Sp[2] + Sp[0] * Tp[1] * Sp[2] + Sp[0] * Sp[1] * Tp[2];
          tm = Tm[0] * Sm[1] * Sm[2] + Sm[0] * Tm[1] * Sm[2] + Sm[0] * Sm[1] * Tm[2];
          const double dS = 2.0 * al * sp_ - ai[k] * sm;
          const double dT = 2.0 * al * tp - ai[k] * tm;
          gA[k] += K * (dmat[a][b] * dT - wmat[a][b] * dS);
        }
      }
    // nuclear attraction: E up to (LA+1, LB+1) (already in Ex..), R up to L+1
    if (c0 < natom) {
      const int c = c0;
      const double X = P[PP_PX] - axyz[3 * c], Y = P[PP_PX + 1] - axyz[3 * c + 1],
                   Z = P[PP_PX + 2] - axyz[3 * c + 2];
      double F[L + 2];
      boys_fn<L + 1, double>(p * (X * X + Y * Y + Z * Z), boys, F);
      double R[L + 2][L + 2][L + 2];
      hermite_r<L + 1, double>(R, F, -2.0 * p, X, Y, Z);
      const double pref = -az[c] * (6.283185307179586477 / p) * K;
      double cA[3] = {0, 0, 0}, cB[3] = {0, 0, 0};
#pragma unroll
      for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const int ai[3] = {CX[LA][a], CY[LA][a], CZ[LA][a]};
          const int bi[3] = {CX[LB][b], CY[LB][b], CZ[LB][b]};
          const double dm = dmat[a][b];
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            // sum_tuv E'_t E_u E_v R_tuv with dimension k differentiated
            double sA = 0.0, sB = 0.0;
#pragma unroll
            for (int t = 0; t <= L + 1; ++t)
#pragma unroll
              for (int u = 0; u <= L + 1 - t; ++u)
#pragma unroll
                for (int v = 0; v <= L + 1 - t - u; ++v) {
                  const int tuv[3] = {t, u, v};
                  double eA = 1.0, eB = 1.0;
                  bool zero = false;
#pragma unroll
                  for (int dd = 0; dd < 3; ++dd) {
                    const double(*E)[LB + 2][L + 4] = dd == 0 ? Ex : (dd == 1 ? Ey : Ez);
                    const int ia = ai[dd], jb = bi[dd], tt = tuv[dd];
                    if (dd == k) {
                      const double da = 2.0 * al * E[ia + 1][jb][tt] - (ia > 0 ? ia * E[ia - 1][jb][tt] : 0.0);
                      const double db = 2.0 * be * E[ia][jb + 1][tt] - (jb > 0 ? jb * E[ia][jb - 1][tt] : 0.0);
                      eA *= da;
                      eB *= db;
                    } else {
                      if (tt > ia + jb) zero = true;
                      const double e0 = tt <= ia + jb ? E[ia][jb][tt] : 0.0;
                      eA *= e0;
                      eB *= e0;
                    }
                  }
                  if (!zero) {
                    sA += eA * R[t][u][v];
                    sB += eB * R[t][u][v];
                  }
                }
            cA[k] += dm * sA;
            cB[k] += dm * sB;
          }
        }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        vA[k] += pref * cA[k];
        vB[k] += pref * cB[k];
        vC[k] -= pref * (cA[k] + cB[k]);
      }
    }
  }
  if (c0 < natom)
#pragma unroll
    for (int k = 0; k < 3; ++k) atomicAdd(&grad[3 * c0 + k], vC[k]);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    atomicAdd(&grad[3 * atA + k], gA[k] + vA[k]);
    atomicAdd(&grad[3 * atB + k], -gA[k] + vB[k]);
  }
}

#define QC_GDECL(A, B, C, D)                                                     \
  hipError_t grad2e_launcht_##A##B##C##D(const Grad2eArgs& a, hipStream_t s);     \
  hipError_t grad2e_launcht32_##A##B##C##D(const Grad2eArgs& a, hipStream_t s);   \
  hipError_t grad2e_launchw_##A##B##C##D(const Grad2eArgs& a, hipStream_t s);   \
  hipError_t grad2e_launchw32_##A##B##C##D(const Grad2eArgs& a, hipStream_t s);  \
  hipError_t grad2e_launchp_##A##B##C##D(const Grad2eArgs& a, hipStream_t s);    \
  hipError_t grad2e_launchp32_##A##B##C##D(const Grad2eArgs& a, hipStream_t s);  \
  hipError_t grad2e_launchwm_##A##B##C##D(const Grad2eArgs& a, hipStream_t s);   \
  hipError_t grad2e_launchwm32_##A##B##C##D(const Grad2eArgs& a, hipStream_t s);  \
  hipError_t grad2e_launchy_##A##B##C##D(const Grad2eArgs& a, hipStream_t s);     \
  hipError_t grad2e_launchy32_##A##B##C##D(const Grad2eArgs& a, hipStream_t s);
QC_ERI_CLASSES(QC_GDECL)
#undef QC_GDECL

}  // namespace qcmd

using namespace qcmd;

// Launch-side density prescreen of one class launch's quartet list: one
// wave per bra pair strides over its kets, keeps those whose bound
//   (Mj[a,b] Mj[c,d] + cx sum_s (Mk_s[a,c] Mk_s[b,d] + Mk_s[a,d] Mk_s[b,c]))
//     Q_ab Q_cd (1 + 1e-6) >= tau
// (shell-block maxima of |Dj|, |Dk_s|; at least the max |G_abcd| the
// gradient kernels' own screen computes, so no quartet they would keep is
// dropped), and either counts them (count mode) or writes them, in list
// order, from the bra's prefix-sum offset (a ballot per 64 kets).
__global__ __launch_bounds__(256) void grad2e_prescreen_kernel(
    int nbra, const int64_t* __restrict__ offsets, const int* __restrict__ bra,
    const int* __restrict__ ket, const int* __restrict__ kstart, const int* __restrict__ shells,
    int ns, const double* __restrict__ Mj, const double* __restrict__ Mk0,
    const double* __restrict__ Mk1, double cx, const double* __restrict__ Q, double tau,
    int64_t* __restrict__ count, const int64_t* __restrict__ write_off, int* __restrict__ out_ket) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (i >= nbra) return;  // whole wave
  const int bp = bra[i];
  const int64_t a = shells[2 * bp], b = shells[2 * bp + 1];
  const double qb = Q[bp] * (1.0 + 1e-6);
  const double mjab = Mj[a * ns + b];
  const int64_t n = offsets[i + 1] - offsets[i];
  const int64_t k0 = kstart[i];
  const int64_t base = write_off ? write_off[i] : 0;
  int64_t kept = 0;
  for (int64_t j0 = 0; j0 < n; j0 += 64) {
    const int64_t j = j0 + lane;
    bool keep = false;
    int kp = 0;
    if (j < n) {
      kp = ket[k0 + j];
      const int64_t c = shells[2 * kp], d = shells[2 * kp + 1];
      double bound = mjab * Mj[c * ns + d];
      bound += cx * (Mk0[a * ns + c] * Mk0[b * ns + d] + Mk0[a * ns + d] * Mk0[b * ns + c]);
      if (Mk1) bound += cx * (Mk1[a * ns + c] * Mk1[b * ns + d] + Mk1[a * ns + d] * Mk1[b * ns + c]);
      keep = bound * qb * Q[kp] >= tau;
    }
    const unsigned long long m = __ballot(keep);
    if (out_ket && keep) out_ket[base + kept + __popcll(m & ((1ull << lane) - 1))] = kp;
    kept += __popcll(m);
  }
  if (!out_ket && lane == 0) count[i] = kept;
}

extern "C" {

int qc_grad1e(int cab, const double* pairs64, const int* sp_prim_off, const int* sp_nprim,
              const int* sp_aoc, const int* sp_atom, const int* sp_diag, const int* pair_list,
              int npairs, int natom, const double* atom_xyz, const double* atom_z, int64_t ncart,
              const double* D, const double* W, double* grad, void* stream) {
  if (npairs <= 0) return 0;
  const double* boys = boys_table64();
  if (!boys) return set_error("qc_grad1e: qc_init not called on the current device");
  if (!D || !W || !grad) return set_error("qc_grad1e: null matrix");
  hipStream_t s = (hipStream_t)stream;
  // pairs over x, nuclei over y (a pair's nuclear loop is the long serial
  // part: few pairs of the high classes, many atoms)
  // and, for classes with few pairs, their primitive pairs over z ((H2O)_16:
  // the 136 (dd) pairs filled 2 x 48 blocks of 128 threads, a tenth of the
  // SIMDs, each thread a spilling serial loop over nine primitive pairs)
  const int pz = npairs < 1024 ? 9 : 1;
  dim3 grid((npairs + 127) / 128, natom < 64 ? (natom > 0 ? natom : 1) : 64, pz), block(128);
#define QC_G1(A, B)                                                                             \
  case (A) * ((A) + 1) / 2 + (B):                                                               \
    hipLaunchKernelGGL((grad1e_kernel<A, B>), grid, block, 0, s, pairs64, sp_prim_off, sp_nprim, \
                       sp_aoc, sp_atom, sp_diag, pair_list, npairs, natom, atom_xyz, atom_z, ncart, \
                       D, W, grad, boys);                                                       \
    break;
  switch (cab) {
    QC_G1(0, 0) QC_G1(1, 0) QC_G1(1, 1) QC_G1(2, 0) QC_G1(2, 1) QC_G1(2, 2)
    default:
      return set_error("qc_grad1e: pair class %d", cab);
  }
#undef QC_G1
  return check(hipGetLastError(), "qc_grad1e launch");
}

int qc_grad2e(int cab, int ccd, const qc_grad2e_args* in, void* stream) {
  if (!in) return set_error("qc_grad2e: null args");
  if (cab < 0 || cab > 5 || ccd < 0 || ccd > cab)
    return set_error("qc_grad2e: pair classes (%d, %d) invalid", cab, ccd);
  if (in->nk < 1 || in->nk > 2) return set_error("qc_grad2e: nk %d", in->nk);
  if (!in->dj || !in->dk0 || (in->nk == 2 && !in->dk1) || !in->grad)
    return set_error("qc_grad2e: null density or gradient");
  Grad2eArgs a;
  a.pairs64 = in->pairs64;
  a.sp_prim_off = in->sp_prim_off;
  a.sp_nprim = in->sp_nprim;
  a.sp_aoc = in->sp_aoc;
  a.sp_atom = in->sp_atom;
  a.sp_diag = in->sp_diag;
  a.bra_list = in->bra_list;
  a.ket_list = in->ket_list;
  a.offsets = in->offsets;
  a.kstart = in->kstart;
  a.nbra = in->nbra;
  a.nquartets = in->nquartets;
  a.ncart = in->ncart;
  a.dj = in->dj;
  a.dk0 = in->dk0;
  a.dk1 = in->dk1;
  a.nk = in->nk;
  a.cx = in->cx;
  a.grad = in->grad;
  if (in->natom < 0) return set_error("qc_grad2e: natom %d", in->natom);
  a.natom = in->natom;
  a.schwarz = in->schwarz;
  a.tau_density = in->tau_density;
  a.pairs32 = in->pairs32;
  a.boys32 = boys_table32();
  if (in->fp32 && !a.pairs32) return set_error("qc_grad2e: fp32 launch without an FP32 pair table");
  if (a.tau_density > 0.0 && !a.schwarz) return set_error("qc_grad2e: tau_density without schwarz");
  a.boys64 = boys_table64();
  if (!a.boys64) return set_error("qc_grad2e: qc_init not called on the current device");
  if (in->variant < 0 || in->variant > 5) return set_error("qc_grad2e: variant %d", in->variant);
  // default (measured on (H2O)_8 and (H2O)_16): one thread per quartet for
  // every class that has that kernel (ket (ss) up to (dp|ss), ket (ps) up to
  // (ds|ps), and (ds|ds)), one wave with the bra side contracted first for
  // the rest
  const bool has_thread = (ccd == 0 && cab <= 4) || (ccd == 1 && cab <= 3) || (cab == 3 && ccd == 3);
  // in FP32 the thread kernel also fits (dd|ss)
  const bool has_thread32 = has_thread || (cab == 5 && ccd == 0);
  // one lane per primitive quartet with G contracted with the bra E first
  // (variant 5) for every class that has it: measured on (H2O)_16 against the
  // previous per-class best (FP64: wave or one lane per primitive quartet
  // without the G-first contraction): (dp|ps) 1074 -> 205 ms, (dp|pp) 1075 ->
  // 296, (dp|dp) 753 -> 164, (dp|ds) 697 -> 116, (ds|pp) 695 -> 248, (pp|pp)
  // 444 -> 157, (dd|ps) 250 -> 74, (dd|pp) 230 -> 94; FP32 (dp|ps) 106 -> 37,
  // (dp|pp) 92 -> 37, (dp|dp) 41 -> 18, (pp|pp) 39 -> 16, (dp|ds) 41 -> 14
  // ((pp|ps) lost: 5.3 -> 10.7 ms FP64, thread form kept)
  // and (dd|ds) 135 -> 35 ms, (dd|dp) 294 -> 91 (FP32 14 -> 4, 30 -> 8),
  // FP64 (dd|ss) 88 -> 20 (FP32 (dd|ss) keeps its thread form)
  const bool yprim = (cab == 2 && ccd == 2) || (cab == 3 && ccd == 2) || (cab == 4 && ccd >= 1) ||
                     (cab == 5 && ccd <= 4);
  const int var = in->variant != 0
                      ? in->variant
                      : ((in->fp32 ? has_thread32 : has_thread) ? 1 : (yprim ? 5 : 2));
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipErrorInvalidValue;
#define QC_GC(A, B, C, D)                                                                     \
  if (cab == (A) * ((A) + 1) / 2 + (B) && ccd == (C) * ((C) + 1) / 2 + (D))                    \
    e = var == 1 ? (in->fp32 ? grad2e_launcht32_##A##B##C##D(a, s) : grad2e_launcht_##A##B##C##D(a, s))  \
      : var == 3 ? (in->fp32 ? grad2e_launchp32_##A##B##C##D(a, s) : grad2e_launchp_##A##B##C##D(a, s))  \
      : var == 4 ? (in->fp32 ? grad2e_launchwm32_##A##B##C##D(a, s) : grad2e_launchwm_##A##B##C##D(a, s)) \
      : var == 5 ? (in->fp32 ? grad2e_launchy32_##A##B##C##D(a, s) : grad2e_launchy_##A##B##C##D(a, s))   \
                 : (in->fp32 ? grad2e_launchw32_##A##B##C##D(a, s) : grad2e_launchw_##A##B##C##D(a, s));
  QC_ERI_CLASSES(QC_GC)
#undef QC_GC
  if (e == hipErrorNotSupported)
    return set_error("qc_grad2e: class (%d, %d) has no %s kernel", cab, ccd,
                     var == 3   ? "one-lane-per-primitive-quartet"
                     : var == 4 ? "matrix-core Y-phase wave"
                     : var == 5 ? "G-first one-lane-per-primitive-quartet"
                                : "one-thread-per-quartet");
  return check(e, "qc_grad2e launch");
}

int qc_grad2e_prescreen(int nbra, const int64_t* offsets, const int* bra, const int* ket,
                        const int* kstart, const int* shells, int ns, const double* Mj,
                        const double* Mk0, const double* Mk1, double cx, const double* Q,
                        double tau, int64_t* count, const int64_t* write_off, int* out_ket,
                        void* stream) {
  if (nbra < 0 || ns <= 0) return set_error("qc_grad2e_prescreen: nbra %d, ns %d", nbra, ns);
  if (!offsets || !bra || !ket || !kstart || !shells || !Mj || !Mk0 || !Q)
    return set_error("qc_grad2e_prescreen: null input");
  if (write_off ? !out_ket : !count)
    return set_error("qc_grad2e_prescreen: need count (count mode) or write_off and out_ket");
  if (nbra == 0) return 0;
  const int per = 4;  // bra pairs (waves) per 256-thread block
  hipLaunchKernelGGL(grad2e_prescreen_kernel, dim3((nbra + per - 1) / per), dim3(64 * per), 0,
                     (hipStream_t)stream, nbra, offsets, bra, ket, kstart, shells, ns, Mj, Mk0,
                     Mk1, cx, Q, tau, count, write_off, out_ket);
  return check(hipGetLastError(), "qc_grad2e_prescreen launch");
}

}  // extern "C"
