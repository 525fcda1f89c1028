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

// ---------------------------------------------------------------------------
// One TC2 step in one launch.  The n x n matrices are cut into 32 x 32 tiles
// (T = ceil(n / 32) per side); one 256-thread work-group per lower-triangle
// tile (tr >= tc) per matrix.  Because X is symmetric, X^2 is too, so only
// these tiles are computed and each off-diagonal one is also written
// transposed.  A work-group first sums the previous launch's per-tile
// partials of X (tr X and |X|_F^2 = tr X^2), takes the TC2 branch from them
// (or leaves a converged matrix unchanged, as qc_tc2_update does), then its
// four waves each compute a 16 x 16 quarter of the tile's X^2 with
// v_mfma_f64_16x16x4f64 (A[i][k] from lane i + 16k, B[k][j] from lane
// j + 16k, output row g + 4r of lane 16g + j in register r), write
// Y = c1 X^2 + c2 X and the tile's own partials of Y for the next step.
// The sum over k is taken in a permuted order, the same for both operands:
// in a block of 16 k, lane (i, kk) supplies k = 4 kk + q to the q-th product,
// so its four values are contiguous (two 16-byte loads per operand).
constexpr int TC2_TILE = 32;

__device__ __forceinline__ void tc2_tile_of(int t, int& tr, int& tc) {
  tr = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while (tr * (tr + 1) / 2 > t) --tr;
  while ((tr + 1) * (tr + 2) / 2 <= t) ++tr;
  tc = t - tr * (tr + 1) / 2;
}

// (sum of the tile's diagonal, sum of squares x (2 off the diagonal)) over
// the 256 threads: each wave reduces with shuffles, wave 0 adds the four.
__device__ __forceinline__ void tc2_block_pair(double a, double c, double* __restrict__ out) {
  __shared__ double s[2][4];
  for (int w = 32; w > 0; w >>= 1) {
    a += __shfl_xor(a, w);
    c += __shfl_xor(c, w);
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s[0][wave] = a;
    s[1][wave] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = s[0][0] + s[0][1] + s[0][2] + s[0][3];
    out[1] = s[1][0] + s[1][1] + s[1][2] + s[1][3];
  }
}

__global__ __launch_bounds__(256) void tc2_tile_partials_kernel(const double* __restrict__ X,
                                                                int64_t n,
                                                                double* __restrict__ part) {
  const int b = blockIdx.y, t = blockIdx.x, ntri = gridDim.x;
  int tr, tc;
  tc2_tile_of(t, tr, tc);
  const double* Xb = X + (int64_t)b * n * n;
  const double w = tr == tc ? 1.0 : 2.0;
  double a = 0.0, c = 0.0;
  for (int e = threadIdx.x; e < TC2_TILE * TC2_TILE; e += 256) {
    const int64_t r = (int64_t)tr * TC2_TILE + e / TC2_TILE, col = (int64_t)tc * TC2_TILE + e % TC2_TILE;
    if (r < n && col < n) {
      const double v = Xb[r * n + col];
      c += w * v * v;
      if (r == col) a += v;
    }
  }
  tc2_block_pair(a, c, part + ((int64_t)b * ntri + t) * 2);
}

__global__ __launch_bounds__(256) void tc2_step_kernel(const double* __restrict__ X,
                                                       double* __restrict__ Y,
                                                       const double* __restrict__ part_in,
                                                       double* __restrict__ part_out,
                                                       const double* __restrict__ nocc,
                                                       int64_t n, double* __restrict__ idem,
                                                       double freeze) {
  typedef double v4 __attribute__((ext_vector_type(4)));
  typedef double v2 __attribute__((ext_vector_type(2)));
  const int b = blockIdx.y, t = blockIdx.x, ntri = gridDim.x;
  int tr, tc;
  tc2_tile_of(t, tr, tc);
  // tr X and |X|_F^2 from the per-tile partials: each thread a strided share,
  // then a fixed-shape tree over the work-group (the same result, and so the
  // same branch, in every work-group of the matrix)
  const double* pb = part_in + (int64_t)b * ntri * 2;
  double trX = 0.0, frX = 0.0;
  for (int k = threadIdx.x; k < ntri; k += 256) {
    trX += pb[2 * k];
    frX += pb[2 * k + 1];
  }
  {
    __shared__ double s[2][4];
    for (int w = 32; w > 0; w >>= 1) {
      trX += __shfl_xor(trX, w);
      frX += __shfl_xor(frX, w);
    }
    if ((threadIdx.x & 63) == 0) {
      s[0][threadIdx.x >> 6] = trX;
      s[1][threadIdx.x >> 6] = frX;
    }
    __syncthreads();
    trX = (s[0][0] + s[0][1]) + (s[0][2] + s[0][3]);
    frX = (s[1][0] + s[1][1]) + (s[1][2] + s[1][3]);
  }
  const double id = fabs(trX - frX);
  if (t == 0 && threadIdx.x == 0) idem[b] = id;
  const double* Xb = X + (int64_t)b * n * n;
  double* Yb = Y + (int64_t)b * n * n;
  const bool frozen = id < freeze;
  const bool square = trX > nocc[b];
  const double c1 = frozen ? 0.0 : (square ? 1.0 : -1.0);
  const double c2 = frozen ? 1.0 : (square ? 0.0 : 2.0);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i16 = lane & 15, g = lane >> 4;
  const int64_t R0 = (int64_t)tr * TC2_TILE, C0 = (int64_t)tc * TC2_TILE;
  // wave w: the whole 32 x 32 tile (2 x 2 blocks of 16 x 16) over its quarter
  // of k (blocks of 16), so every loaded fragment feeds two products
  const int64_t kq = ((n + 63) / 64) * 16;
  const int64_t kbeg = kq * wave, kend = kq * (wave + 1) < n ? kq * (wave + 1) : n;
  v4 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) acc[x][y] = (v4){0.0, 0.0, 0.0, 0.0};
  if (!frozen) {
    const double* pr[2];
    const double* pc[2];
    bool inr[2], inc[2];
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      const int64_t ra = R0 + 16 * x + i16, rb = C0 + 16 * x + i16;
      inr[x] = ra < n;
      inc[x] = rb < n;
      pr[x] = Xb + (inr[x] ? ra : 0) * n;
      pc[x] = Xb + (inc[x] ? rb : 0) * n;
    }
    const bool vec = (n & 1) == 0;  // 16-byte aligned rows
#pragma unroll 2
    for (int64_t k0 = kbeg; k0 < kend; k0 += 16) {
      const int64_t k = k0 + 4 * g;
      double a[2][4], bv[2][4];
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        if (vec && k + 4 <= kend) {
          const v2 a0 = *reinterpret_cast<const v2*>(pr[x] + k), a1 = *reinterpret_cast<const v2*>(pr[x] + k + 2);
          const v2 b0 = *reinterpret_cast<const v2*>(pc[x] + k), b1 = *reinterpret_cast<const v2*>(pc[x] + k + 2);
          a[x][0] = a0.x; a[x][1] = a0.y; a[x][2] = a1.x; a[x][3] = a1.y;
          bv[x][0] = b0.x; bv[x][1] = b0.y; bv[x][2] = b1.x; bv[x][3] = b1.y;
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            a[x][q] = k + q < kend ? pr[x][k + q] : 0.0;
            bv[x][q] = k + q < kend ? pc[x][k + q] : 0.0;
          }
        }
        if (!inr[x]) a[x][0] = a[x][1] = a[x][2] = a[x][3] = 0.0;
        if (!inc[x]) bv[x][0] = bv[x][1] = bv[x][2] = bv[x][3] = 0.0;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
          for (int y = 0; y < 2; ++y)
            acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x][q], bv[y][q], acc[x][y], 0, 0, 0);
    }
  }
  // the four waves' partial tiles through LDS, summed in wave order
  __shared__ double red[4][TC2_TILE][TC2_TILE + 1];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) red[wave][16 * x + g + 4 * r][16 * y + i16] = acc[x][y][r];
  __syncthreads();
  const bool diag = tr == tc;
  double sa = 0.0, sc = 0.0;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int e = threadIdx.x + 256 * m, lr = e / TC2_TILE, lc = e % TC2_TILE;
    const int64_t row = R0 + lr, col = C0 + lc;
    if (row < n && col < n) {
      const double p2 = ((red[0][lr][lc] + red[1][lr][lc]) + red[2][lr][lc]) + red[3][lr][lc];
      const double y = c1 * p2 + c2 * Xb[row * n + col];
      Yb[row * n + col] = y;
      if (!diag) Yb[col * n + row] = y;
      sc += (diag ? 1.0 : 2.0) * y * y;
      if (row == col) sa += y;
    }
  }
  tc2_block_pair(sa, sc, part_out + ((int64_t)b * ntri + t) * 2);
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

static int tc2_ntri(int64_t n) {
  const int64_t T = (n + TC2_TILE - 1) / TC2_TILE;
  return (int)(T * (T + 1) / 2);
}

int64_t qc_tc2_tile_partials_len(int64_t n, int nb) { return (int64_t)nb * tc2_ntri(n) * 2; }

int qc_tc2_tile_partials(const double* X, int64_t n, int nb, double* part, void* stream) {
  if (n <= 0 || n > 32768 || nb <= 0 || nb > 65535)
    return set_error("qc_tc2_tile_partials: n %lld, nb %d", (long long)n, nb);
  if (!X || !part) return set_error("qc_tc2_tile_partials: null pointer");
  hipLaunchKernelGGL(tc2_tile_partials_kernel, dim3(tc2_ntri(n), nb), dim3(256), 0,
                     (hipStream_t)stream, X, n, part);
  return check(hipGetLastError(), "qc_tc2_tile_partials");
}

int qc_tc2_step(const double* X, double* Y, const double* part_in, double* part_out,
                const double* nocc, int64_t n, int nb, double* idem, double freeze, void* stream) {
  if (n <= 0 || n > 32768 || nb <= 0 || nb > 65535)
    return set_error("qc_tc2_step: n %lld, nb %d", (long long)n, nb);
  if (!X || !Y || !part_in || !part_out || !nocc || !idem) return set_error("qc_tc2_step: null pointer");
  if (X == Y || part_in == part_out) return set_error("qc_tc2_step: in-place step is not supported");
  hipLaunchKernelGGL(tc2_step_kernel, dim3(tc2_ntri(n), nb), dim3(256), 0, (hipStream_t)stream, X, Y,
                     part_in, part_out, nocc, n, idem, freeze);
  return check(hipGetLastError(), "qc_tc2_step");
}

}  // extern "C"
