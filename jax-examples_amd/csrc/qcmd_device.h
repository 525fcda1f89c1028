// qcmd_device.h -- device-side building blocks of the McMurchie-Davidson path
// for gfx950: Boys function, 1-D Hermite expansion coefficients, the R_tuv
// recursion and Cartesian/spherical component tables.  Everything indexed by
// angular momentum is resolved at compile time so that the per-class kernels
// keep their intermediates in VGPRs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qcmd {

// ---------------------------------------------------------------------------
// Boys function F_n(T) = int_0^1 u^{2n} exp(-T u^2) du.
//
// Method: pretabulated Taylor expansion with downward recursion (Gill, Johnson
// & Pople, Int. J. Quantum Chem. 40, 745 (1991); Helgaker, Jorgensen & Olsen,
// "Molecular Electronic-Structure Theory", sec. 9.8.2):
//   F_N(T) = sum_k F_{N+k}(T_i) (-d)^k / k!,  d = T - T_i, |d| <= h/2,
//   F_{n}(T) = (2T F_{n+1}(T) + exp(-T)) / (2n+1)   (stable downward)
// on a grid T_i = i*h for T < BOYS_TMAX.  Above BOYS_TMAX, F_0 is the erf
// closed form (erfc(sqrt T) below 1e-19 relative) followed by the upward
// recursion, stable there because T > n.
// ---------------------------------------------------------------------------
constexpr int BOYS_NMAX = 12;                 // highest n served (derivative ERIs need L+1 <= 9)
constexpr int BOYS_TAYLOR64 = 7;              // terms for FP64 (error ~ (h/2)^7/7! ~ 1e-15)
constexpr int BOYS_TAYLOR32 = 4;              // terms for FP32 (error ~ 1.6e-8)
constexpr int BOYS_NCOL = BOYS_NMAX + BOYS_TAYLOR64;  // table columns F_0..F_{NCOL-1}
constexpr double BOYS_H = 0.05;
constexpr double BOYS_TMAX = 40.0;
constexpr int BOYS_NGRID = 801;               // T_i = 0 .. 40 inclusive

template <typename Real>
struct BoysTable;
template <>
struct BoysTable<double> {
  const double* t;
};
template <>
struct BoysTable<float> {
  const float* t;
};

__device__ inline double dev_exp(double x) { return exp(x); }
__device__ inline float dev_exp(float x) { return __expf(x); }
__device__ inline double dev_sqrt(double x) { return sqrt(x); }
__device__ inline float dev_sqrt(float x) { return sqrtf(x); }
__device__ inline double dev_rcp(double x) { return 1.0 / x; }
__device__ inline float dev_rcp(float x) { return __frcp_rn(x); }

// F[0..N] at T.  tab: [BOYS_NGRID][BOYS_NCOL] row-major.
template <int N, typename Real>
__device__ __forceinline__ void boys_fn(Real T, const Real* __restrict__ tab, Real (&F)[N + 1]) {
  constexpr int K = sizeof(Real) == 8 ? BOYS_TAYLOR64 : BOYS_TAYLOR32;
  if (T < (Real)BOYS_TMAX) {
    int i = (int)(T * (Real)(1.0 / BOYS_H) + (Real)0.5);
    Real d = (Real)(i * BOYS_H) - T;  // -(T - T_i)
    const Real* row = tab + i * BOYS_NCOL + N;
    // Horner in d: sum_k F_{N+k} d^k / k!
    Real f = row[K - 1];
#pragma unroll
    for (int k = K - 1; k >= 1; --k) f = row[k - 1] + f * d * (Real)(1.0 / k);
    F[N] = f;
    if (N > 0) {
      Real e = dev_exp(-T);
      Real t2 = T + T;
#pragma unroll
      for (int n = N - 1; n >= 0; --n) F[n] = (t2 * F[n + 1] + e) * (Real)(1.0 / (2 * n + 1));
    }
  } else {
    Real rt = dev_rcp(T);
    F[0] = (Real)0.88622692545275801365 * dev_sqrt(rt);  // sqrt(pi)/2 / sqrt(T)
    if (N > 0) {
      Real e = dev_exp(-T);
      Real half_rt = (Real)0.5 * rt;
#pragma unroll
      for (int n = 0; n < N; ++n) F[n + 1] = ((Real)(2 * n + 1) * F[n] - e) * half_rt;
    }
  }
}

// F[0..N] at T for a kernel computing in Real.  FP64 kernels evaluate the
// FP64 grid directly.  FP32 kernels (the mixed path's primitive sums)
// evaluate the same FP64 grid, Taylor series and recursion in FP64 and round
// the values to FP32: the Boys function stays an FP64 evaluation, only the
// primitive sums that consume it run in FP32.
template <int N, typename Real>
__device__ __forceinline__ void boys_eval(Real T, const double* __restrict__ tab64,
                                          Real (&F)[N + 1]) {
  if constexpr (sizeof(Real) == 8) {
    boys_fn<N, double>(T, tab64, F);
  } else {
    double Fd[N + 1];
    boys_fn<N, double>((double)T, tab64, Fd);
#pragma unroll
    for (int n = 0; n <= N; ++n) F[n] = (Real)Fd[n];
  }
}

// ---------------------------------------------------------------------------
// Component tables
// ---------------------------------------------------------------------------
// Launch shape for nblocks work-groups of `block` lanes.  A dispatch packet
// holds the grid size of each dimension as a 32-bit work-item count, so a
// launch whose x extent would pass 2^31 work-items spills into y; kernels
// take the flat work-group index from flat_block() and bound it themselves.
inline dim3 grid_1d(int64_t nblocks, int block) {
  const int64_t maxx = (int64_t(1) << 31) / block;
  if (nblocks <= maxx) return dim3((unsigned)(nblocks > 0 ? nblocks : 1));
  return dim3((unsigned)maxx, (unsigned)((nblocks + maxx - 1) / maxx));
}
__device__ __forceinline__ int64_t flat_block() {
  return (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
}

__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ constexpr int nsph(int l) { return 2 * l + 1; }
__host__ __device__ constexpr int nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }

// Cartesian component k of shell l, ordered lx descending then ly descending
// (jax_qc/integrals/mcmurchie_davidson.py:39-50).
__host__ __device__ constexpr int cart_lx(int l, int k) {
  int idx = 0;
  for (int lx = l; lx >= 0; --lx)
    for (int ly = l - lx; ly >= 0; --ly) {
      if (idx == k) return lx;
      ++idx;
    }
  return 0;
}
__host__ __device__ constexpr int cart_ly(int l, int k) {
  int idx = 0;
  for (int lx = l; lx >= 0; --lx)
    for (int ly = l - lx; ly >= 0; --ly) {
      if (idx == k) return ly;
      ++idx;
    }
  return 0;
}
__host__ __device__ constexpr int cart_lz(int l, int k) { return l - cart_lx(l, k) - cart_ly(l, k); }

// Cartesian -> real spherical coefficient (row s, Cartesian column k) in the
// libcint row order of jax_qc/integrals/spherical.py:34-67.  l = 1 is the
// identity (px, py, pz).
__host__ __device__ constexpr double sph_coef(int l, int s, int k) {
  if (l == 0) return 1.0;
  if (l == 1) return s == k ? 1.0 : 0.0;
  // l == 2: columns xx, xy, xz, yy, yz, zz
  constexpr double r3 = 1.7320508075688772935;
  switch (s) {
    case 0: return k == 1 ? r3 : 0.0;                              // d_xy
    case 1: return k == 4 ? r3 : 0.0;                              // d_yz
    case 2: return k == 0 ? -0.5 : (k == 3 ? -0.5 : (k == 5 ? 1.0 : 0.0));  // d_z2
    case 3: return k == 2 ? r3 : 0.0;                              // d_xz
    default: return k == 0 ? 0.5 * r3 : (k == 3 ? -0.5 * r3 : 0.0);      // d_x2-y2
  }
}

// ---------------------------------------------------------------------------
// 1-D Hermite expansion coefficients E[i][j][t], i<=LA, j<=LB, t<=i+j, with
// E[0][0][0] = 1 (the K_AB factor is carried separately).  Same recurrence
// and increment order as jax_qc/integrals/mcmurchie_davidson.py:60-108.
// ---------------------------------------------------------------------------
template <int LA, int LB, typename Real>
__device__ __forceinline__ void hermite_e(Real (&E)[LA + 1][LB + 1][LA + LB + 1], Real PA, Real PB,
                                          Real oo2p) {
#pragma unroll
  for (int i = 0; i <= LA; ++i)
#pragma unroll
    for (int j = 0; j <= LB; ++j)
#pragma unroll
      for (int t = 0; t <= LA + LB; ++t) E[i][j][t] = (Real)0;
  E[0][0][0] = (Real)1;
#pragma unroll
  for (int i = 0; i < LA; ++i) {
#pragma unroll
    for (int t = 0; t <= i + 1; ++t) {
      Real v = PA * (t <= i ? E[i][0][t] : (Real)0);
      if (t >= 1) v += oo2p * E[i][0][t - 1];
      if (t + 1 <= i) v += (Real)(t + 1) * E[i][0][t + 1];
      E[i + 1][0][t] = v;
    }
  }
#pragma unroll
  for (int j = 0; j < LB; ++j)
#pragma unroll
    for (int i = 0; i <= LA; ++i)
#pragma unroll
      for (int t = 0; t <= i + j + 1; ++t) {
        Real v = PB * (t <= i + j ? E[i][j][t] : (Real)0);
        if (t >= 1) v += oo2p * E[i][j][t - 1];
        if (t + 1 <= i + j) v += (Real)(t + 1) * E[i][j][t + 1];
        E[i][j + 1][t] = v;
      }
}

// ---------------------------------------------------------------------------
// R^0_{tuv}(alpha, X, Y, Z) for t+u+v <= L, from F_0..F_L by the recursion of
// jax_qc/integrals/mcmurchie_davidson.py:113-172, run level by level from
// n = L down to 0 and only over the tetrahedron t+u+v <= L-n:
//   R^n_{000} = (-2 alpha)^n F_n
//   R^n_{t+1,u,v} = t R^{n+1}_{t-1,u,v} + X R^{n+1}_{t,u,v}   (then u, then v)
// R is indexed [t][u][v]; entries with t+u+v > L are left unset.
// ---------------------------------------------------------------------------
template <int L, typename Real>
__device__ __forceinline__ void hermite_r(Real (&R)[L + 1][L + 1][L + 1], const Real (&F)[L + 1],
                                          Real m2a, Real X, Real Y, Real Z) {
  Real P[L + 1][L + 1][L + 1];
  Real pw = (Real)1;
#pragma unroll
  for (int n = 0; n < L; ++n) pw *= m2a;
  P[0][0][0] = pw * F[L];
#pragma unroll
  for (int n = L - 1; n >= 0; --n) {
    pw = (Real)1;
#pragma unroll
    for (int k = 0; k < n; ++k) pw *= m2a;
    Real C[L + 1][L + 1][L + 1];
    C[0][0][0] = pw * F[n];
    constexpr int dummy = 0;
    (void)dummy;
#pragma unroll
    for (int t = 0; t <= L - n; ++t)
#pragma unroll
      for (int u = 0; u <= L - n - t; ++u)
#pragma unroll
        for (int v = 0; v <= L - n - t - u; ++v) {
          if (t == 0 && u == 0 && v == 0) continue;
          Real r;
          if (t > 0) {
            r = X * P[t - 1][u][v];
            if (t > 1) r += (Real)(t - 1) * P[t - 2][u][v];
          } else if (u > 0) {
            r = Y * P[t][u - 1][v];
            if (u > 1) r += (Real)(u - 1) * P[t][u - 2][v];
          } else {
            r = Z * P[t][u][v - 1];
            if (v > 1) r += (Real)(v - 1) * P[t][u][v - 2];
          }
          C[t][u][v] = r;
        }
#pragma unroll
    for (int t = 0; t <= L - n; ++t)
#pragma unroll
      for (int u = 0; u <= L - n - t; ++u)
#pragma unroll
        for (int v = 0; v <= L - n - t - u; ++v) P[t][u][v] = C[t][u][v];
  }
#pragma unroll
  for (int t = 0; t <= L; ++t)
#pragma unroll
    for (int u = 0; u <= L - t; ++u)
#pragma unroll
      for (int v = 0; v <= L - t - u; ++v) R[t][u][v] = P[t][u][v];
}

// Packed-pair helpers for basis-function pair indices.
__host__ __device__ inline int64_t pair_index(int64_t i, int64_t j) {
  return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
}

}  // namespace qcmd
