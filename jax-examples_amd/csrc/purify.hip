// purify.hip -- the non-GEMM parts of one trace-correcting purification
// step (TC2, A. M. N. Niklasson, Phys. Rev. B 66, 155115 (2002)) for a batch
// of symmetric n x n matrices X_b:
//
//   X_b <- X_b^2            if tr X_b > n_occ[b]
//   X_b <- 2 X_b - X_b^2    otherwise
//
// The SCF drivers need the projector onto the occupied subspace of the Fock
// matrix at every step (the reference diagonalises instead, jax_qc/scf/
// rhf.py:52-92, uhf.py:187-287).  A step is qc_tc2_partials (tr X and
// |X|_F^2 = tr X^2 for symmetric X, as per-block partial sums), the product
// X^2 (a library GEMM), and qc_tc2_update (the branch taken on the device
// from the partials and the new X written in place), so no host
// synchronisation and no reduction library call sits inside the loop and a
// chunk of steps can be captured in one graph.
#include "qcmd_device.h"
#include "qcmd_internal.h"

namespace qcmd {

constexpr int TC2_BLOCKS = 64;   // partial-sum blocks per matrix
constexpr int TC2_THREADS = 256;

// part[b][blk] = (sum of X_ii, sum of X_ij^2) over block blk's share of
// matrix b (elements and diagonal entries blk * TC2_THREADS + thread, then
// strides of TC2_BLOCKS * TC2_THREADS).
__global__ __launch_bounds__(TC2_THREADS) void tc2_partials_kernel(const double* __restrict__ X,
                                                                   int64_t n,
                                                                   double* __restrict__ part) {
  const int b = blockIdx.y, blk = blockIdx.x;
  const double* Xb = X + (int64_t)b * n * n;
  const int64_t nn = n * n;
  double tr = 0.0, fr = 0.0;
  for (int64_t i = (int64_t)blk * TC2_THREADS + threadIdx.x; i < nn;
       i += (int64_t)TC2_BLOCKS * TC2_THREADS) {
    const double v = Xb[i];
    fr += v * v;
  }
  for (int64_t d = (int64_t)blk * TC2_THREADS + threadIdx.x; d < n;
       d += (int64_t)TC2_BLOCKS * TC2_THREADS)
    tr += Xb[d * (n + 1)];
  for (int w = 32; w > 0; w >>= 1) {
    tr += __shfl_xor(tr, w);
    fr += __shfl_xor(fr, w);
  }
  __shared__ double s[2][TC2_THREADS / 64];
  if ((threadIdx.x & 63) == 0) {
    s[0][threadIdx.x >> 6] = tr;
    s[1][threadIdx.x >> 6] = fr;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, c = 0.0;
#pragma unroll
    for (int w = 0; w < TC2_THREADS / 64; ++w) {
      a += s[0][w];
      c += s[1][w];
    }
    part[((int64_t)b * TC2_BLOCKS + blk) * 2] = a;
    part[((int64_t)b * TC2_BLOCKS + blk) * 2 + 1] = c;
  }
}

// Every block sums matrix b's partials itself (TC2_BLOCKS pairs, in a fixed
// order, so every block takes the same branch), then writes its share of
// X_b <- c1 X2_b + c2 X_b, unless |tr X_b - tr X_b^2| is already below
// freeze: a converged X is left as it is (further products would amplify
// its rounding -- an eigenvalue 1 + e becomes 1 + 2e under X^2, -e becomes
// -2e under 2X - X^2 -- and diverge), so steps replayed past convergence are
// harmless.  Block 0 also stores tr X_b and idem[b] = |tr X_b - tr X_b^2|
// (the distance from idempotency) of the matrix before the step.
__global__ __launch_bounds__(TC2_THREADS) void tc2_update_kernel(double* __restrict__ X,
                                                                 const double* __restrict__ X2,
                                                                 const double* __restrict__ part,
                                                                 const double* __restrict__ nocc,
                                                                 int64_t n, double* __restrict__ idem,
                                                                 double* __restrict__ trace,
                                                                 double freeze) {
  const int b = blockIdx.y, blk = blockIdx.x;
  double tr = 0.0, fr = 0.0;
  for (int k = 0; k < TC2_BLOCKS; ++k) {
    tr += part[((int64_t)b * TC2_BLOCKS + k) * 2];
    fr += part[((int64_t)b * TC2_BLOCKS + k) * 2 + 1];
  }
  const bool square = tr > nocc[b];
  const double c1 = square ? 1.0 : -1.0, c2 = square ? 0.0 : 2.0;
  double* Xb = X + (int64_t)b * n * n;
  const double* X2b = X2 + (int64_t)b * n * n;
  const int64_t nn = n * n;
  if (!(fabs(tr - fr) < freeze))
    for (int64_t i = (int64_t)blk * TC2_THREADS + threadIdx.x; i < nn;
         i += (int64_t)TC2_BLOCKS * TC2_THREADS)
      Xb[i] = c1 * X2b[i] + c2 * Xb[i];
  if (blk == 0 && threadIdx.x == 0) {
    idem[b] = fabs(tr - fr);
    if (trace) trace[b] = tr;
  }
}

}  // namespace qcmd

using namespace qcmd;

extern "C" {

int64_t qc_tc2_partials_len(int nb) { return (int64_t)nb * TC2_BLOCKS * 2; }

int qc_tc2_partials(const double* X, int64_t n, int nb, double* part, void* stream) {
  if (n <= 0 || nb <= 0) return set_error("qc_tc2_partials: n %lld, nb %d", (long long)n, nb);
  if (nb > 65535) return set_error("qc_tc2_partials: nb %d above 65535", nb);
  if (!X || !part) return set_error("qc_tc2_partials: null pointer");
  hipLaunchKernelGGL(tc2_partials_kernel, dim3(TC2_BLOCKS, nb), dim3(TC2_THREADS), 0,
                     (hipStream_t)stream, X, n, part);
  return check(hipGetLastError(), "qc_tc2_partials");
}

int qc_tc2_update(double* X, const double* X2, const double* part, const double* nocc, int64_t n,
                  int nb, double* idem, double* trace, double freeze, void* stream) {
  if (n <= 0 || nb <= 0) return set_error("qc_tc2_update: n %lld, nb %d", (long long)n, nb);
  if (nb > 65535) return set_error("qc_tc2_update: nb %d above 65535", nb);
  if (!X || !X2 || !part || !nocc || !idem) return set_error("qc_tc2_update: null pointer");
  hipLaunchKernelGGL(tc2_update_kernel, dim3(TC2_BLOCKS, nb), dim3(TC2_THREADS), 0,
                     (hipStream_t)stream, X, X2, part, nocc, n, idem, trace, freeze);
  return check(hipGetLastError(), "qc_tc2_update");
}

}  // extern "C"
