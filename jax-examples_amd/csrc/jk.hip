// jk.hip -- Fock build from integrals held in HBM.
//
// The reference contracts a dense (n,n,n,n) tensor with two einsums per Fock
// build (jax_qc/scf/fock.py:31-79).  Here the unique integrals live in a
// packed supermatrix over basis-function pairs p = i(i+1)/2 + j (i >= j),
// tile-major in 64 x 64 tiles (lower triangle of tiles, diagonal tiles
// full).  With D~_kl = D_kl (2 - delta_kl):
//   J_ij = sum_{k>=l} (ij|kl) D~_kl
//   K_ij = sum_{k>=l} 1/2[(ik|jl) + (il|jk)] D~_kl
//   RHF: F = H + sum_{k>=l} P[(ij),(kl)] D~_kl,  P = (ij|kl) - 1/4[(ik|jl)+(il|jk)]
// (Raffenetti's supermatrix), so each Fock build is one symmetric
// matrix-vector product that streams every unique integral once from HBM.
#include <mutex>
#include <type_traits>

#include "qcmd_device.h"
#include "qcmd_internal.h"
#include "qcmd_types.h"

namespace qcmd {

// Tiles per SYMV work item and the load flavour; set by qc_symv_config.
// 16 tiles per item measured best on the (H2O)_16 pools (K pass with two
// right-hand sides: 6.14 ms at 4, 5.29 at 8, 5.04 at 16, 4.96 at 32; the
// one-right-hand-side J pass 5.38 / 5.31 / 5.31 / 5.35 ms).
static int g_symv_seg = 16;
static int g_symv_nt = 1;
// Waves per work item by right-hand-side count (qc_symv_split): with one
// wave a lane holds 16 rows of x and row sums per right-hand side, which at
// two right-hand sides costs the second wave per SIMD (334 VGPRs); two waves
// per item halve that (198 VGPRs; K pass 5.40 -> 5.04 ms at 16 tiles per
// item).  One right-hand side measured the same either way.
static int g_symv_ws[5] = {1, 1, 2, 2, 2};

// Workspace of qc_symv_noalloc: the SYMV's row and column partials, for
// callers that cannot allocate (graph capture, serving loops).  Each GPU
// holds its own copy for the life of the process; its address is taken per
// device at qc_init (host code never uses the symbol itself).  Every call on
// a GPU shares that GPU's copy, so calls that overlap on one GPU race.
constexpr int64_t kSymvScratchBytes = int64_t(64) << 20;
static __device__ __align__(256) double scratch[kSymvScratchBytes / sizeof(double)];
constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

static void* symv_scratch_ptr(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[dev];
}
typedef double dvec2 __attribute__((ext_vector_type(2)));

__host__ __device__ inline int64_t sm_tiles_per_dim(int64_t nbf) {
  int64_t npair = nbf * (nbf + 1) / 2;
  return (npair + SM_B - 1) / SM_B;
}

__device__ inline int64_t tri_root(int64_t t) {  // largest R with R(R+1)/2 <= t
  int64_t R = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while (R * (R + 1) / 2 > t) --R;
  while ((R + 1) * (R + 2) / 2 <= t) ++R;
  return R;
}

__device__ inline int64_t sm_pos2(int64_t p, int64_t q) {  // any order
  if (p < q) {
    int64_t t = p;
    p = q;
    q = t;
  }
  int64_t R = p >> SM_LOGB, C = q >> SM_LOGB;
  return ((R * (R + 1) / 2 + C) << (2 * SM_LOGB)) + ((p & (SM_B - 1)) << SM_LOGB) + (q & (SM_B - 1));
}

__device__ inline int64_t pidx(int64_t i, int64_t j) {
  return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
}

// Max of m over a 256-lane work-group, written to *dst by lane 0 (every
// lane of the work-group must call it).
__device__ __forceinline__ void tile_block_max(double m, double* dst) {
  __shared__ double wmax[4];
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) *dst = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
}

// One work-group per 64 x 64 output tile (R, C); lane c of wave w computes
// the elements (r, c), r = w, w + 4, ..., w + 60, so each wave writes whole
// tile rows.  Row r is the pair p = (ij), column c the pair q = (kl), and
// K_pq = [(ik|jl) + (il|jk)] / 2 (both orders of p and q give the same
// value, so diagonal tiles hold both triangles).  Over the tile the first
// integral reads row (ik) of J at the 64 x 64 columns (jl), and the second
// the 64 x 64 block of rows (il), columns (jk), both gathered element by
// element; where j > k the second has no contiguous direction in the tile
// (one cache line per element), which bounds the kernel: (H2O)_16, 53 GB
// per matrix, 63 ms (68 ms with a work-group per four tile rows).
__global__ __launch_bounds__(256) void sm_assemble_kernel(int kind, const double* __restrict__ J,
                                                          double* __restrict__ out, int64_t nbf,
                                                          int64_t ntiles, double* __restrict__ tmax) {
  const int64_t t = flat_block();
  if (t >= ntiles) return;
  const int64_t npair = nbf * (nbf + 1) / 2;
  const int64_t R = tri_root(t), C = t - R * (R + 1) / 2;
  const int c = threadIdx.x & (SM_B - 1), w = threadIdx.x >> SM_LOGB;
  const int64_t q = C * SM_B + c;
  const bool qok = q < npair;
  const int64_t k = tri_root(q), l = q - k * (k + 1) / 2;
  int64_t p = R * SM_B + w;
  int64_t i = tri_root(p), j = p - i * (i + 1) / 2;
  double* o = out + (t << (2 * SM_LOGB)) + c;
  double m = 0.0;
  for (int r = w; r < SM_B; r += 4) {
    double v = 0.0;
    if (qok && p < npair) {
      const double v1 = J[sm_pos2(pidx(i, k), pidx(j, l))];
      const double v2 = J[sm_pos2(pidx(i, l), pidx(j, k))];
      const double kv = 0.5 * (v1 + v2);
      v = kind == 0 ? J[sm_pos2(p, q)] - 0.5 * kv : kv;
    }
    __builtin_nontemporal_store(v, o + r * SM_B);
    m = fmax(m, fabs(v));
    p += 4;  // next row of this wave: (ij) four pairs on
    j += 4;
    while (j > i) {
      j -= i + 1;
      ++i;
    }
  }
  if (tmax) tile_block_max(m, tmax + t);
}

// J <- J - K/2 in place, one work-group per 64 x 64 tile (both tile-major
// supermatrices; the padding is zero in both): a lane takes eight pairs of
// doubles, all sixteen loads issued before the stores, and with tmax the
// tile's max |J - K/2| is written to tmax[tile] (the compaction's tile
// maxima without another read of the matrix).
__global__ __launch_bounds__(256) void sm_p_from_jk_kernel(double* __restrict__ J,
                                                           const double* __restrict__ K,
                                                           int64_t ntiles,
                                                           double* __restrict__ tmax) {
  const int64_t t = flat_block();
  if (t >= ntiles) return;
  dvec2* j2 = reinterpret_cast<dvec2*>(J) + (t << (2 * SM_LOGB - 1)) + threadIdx.x;
  const dvec2* k2 = reinterpret_cast<const dvec2*>(K) + (t << (2 * SM_LOGB - 1)) + threadIdx.x;
  dvec2 a[8], b[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    a[u] = __builtin_nontemporal_load(j2 + 256 * u);
    b[u] = __builtin_nontemporal_load(k2 + 256 * u);
  }
  double m = 0.0;
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const dvec2 v = a[u] - 0.5 * b[u];
    __builtin_nontemporal_store(v, j2 + 256 * u);
    m = fmax(m, fmax(fabs(v.x), fabs(v.y)));
  }
  if (tmax) tile_block_max(m, tmax + t);
}

// Work item m of the SYMV: tile row R, segment s covering tiles
// [s*seg, min(s*seg+seg, R+1)).  Row R has R/seg + 1 segments, so
// item_off(R) = R + seg*m(m-1)/2 + m*(R - m*seg) with m = R/seg.
__host__ __device__ inline int64_t item_off(int64_t R, int seg) {
  const int64_t m = R / seg;
  return R + seg * m * (m - 1) / 2 + m * (R - m * seg);
}

// Bytes of row and column partials one SYMV uses at segment length seg:
// item_off(T, seg) row-partial blocks and T(T+1)/2 column-partial blocks of
// 64 doubles per right-hand side.
static int64_t symv_work_bytes(int64_t nbf, int nrhs, int seg) {
  const int64_t T = sm_tiles_per_dim(nbf);
  const int64_t items = item_off(T, seg), tiles = T * (T + 1) / 2;
  return (items + tiles) * nrhs * SM_B * (int64_t)sizeof(double);
}

// Tile formats of a compacted (mixed-precision) supermatrix.  The two
// integer formats hold q = rint(x / s) per element with one FP64 scale s =
// max|x| / 32767 (int16) or max|x| / 127 (int8) per tile, so an element is
// off by at most s / 2 = max|x| / 65534 or max|x| / 254: a fixed fraction of
// the tile's largest element (an FP32 tile's error is relative to each
// element instead).  Their pool block is a 16-byte header (s, then padding)
// followed by the 64 x 64 integers, row-major.
enum TileFmt : uint8_t { TILE_ZERO = 0, TILE_F32 = 1, TILE_F64 = 2, TILE_I16 = 3, TILE_I8 = 4 };
constexpr int TILE_QHDR = 16;  // bytes before the integers of an int16 / int8 tile
typedef float fvec2 __attribute__((ext_vector_type(2)));
typedef int ivec2 __attribute__((ext_vector_type(2)));

typedef float fvec4 __attribute__((ext_vector_type(4)));

template <bool NT, typename V>
__device__ __forceinline__ V ld(const V* p) {
  if (NT) return __builtin_nontemporal_load(p);
  return *p;
}

// Lane layout of the tile pass: lane = (ro, cg), ro = lane >> 4 in 0..3,
// cg = lane & 15; the lane covers columns 4cg..4cg+3 of every tile and rows
// 4k + ro, k < 16.  Both tile formats are row-major, so a wave reads 1 KB
// (FP32) or 2 KB (FP64) of contiguous tile per row quadruple.
constexpr int SYMV_ROWS = 16;  // rows per lane
#ifndef SYMV_MIN_WAVES
#define SYMV_MIN_WAVES 1  // waves per SIMD the register budget must allow
#endif

// Column j < 4 of lane cg: FP32 / int16 / int8 tiles 4cg + j (16 / 8 / 4 B
// per row); FP64 tiles 2cg + j for j < 2 and 32 + 2cg + (j-2) otherwise, so
// each 16-B load of a wave covers 256 contiguous bytes of four rows.
template <typename T>
__device__ __forceinline__ int lane_col(int cg, int j) {
  if (sizeof(T) != 8) return 4 * cg + j;
  return j < 2 ? 2 * cg + j : 32 + 2 * cg + (j - 2);
}

// Column partials of one off-diagonal tile: sum over the lane's rows, then
// over the four row offsets (xor 16, xor 32); lanes with ro == 0 store.  With
// two waves per work item (WS = 2, each wave half of the tile's rows) wave 1
// hands its sums to wave 0 through LDS, double-buffered by the tile's parity
// so that one barrier per tile suffices: wave 1 writes buffer p for tile t+2
// only after both waves passed tile t+1's barrier, by which time wave 0 has
// read tile t's values from it.
template <int NRHS, int WS>
struct ColShare {
  double v[2][NRHS][SM_B];
};
template <int NRHS>
struct ColShare<NRHS, 1> {};

template <int NRHS, int WS, typename T>
__device__ __forceinline__ void store_cols(double (&col)[NRHS][4], int64_t tile, int cg, int ro,
                                           int wave, int par, ColShare<NRHS, WS>& sh,
                                           double* __restrict__ colbuf) {
#pragma unroll
  for (int h = 0; h < NRHS; ++h)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      col[h][j] += __shfl_xor(col[h][j], 16);
      col[h][j] += __shfl_xor(col[h][j], 32);
    }
  if constexpr (WS == 2) {
    if (wave == 1 && ro == 0) {
#pragma unroll
      for (int h = 0; h < NRHS; ++h)
#pragma unroll
        for (int j = 0; j < 4; ++j) sh.v[par][h][lane_col<T>(cg, j)] = col[h][j];
    }
    __syncthreads();
    if (wave == 1) return;
#pragma unroll
    for (int h = 0; h < NRHS; ++h)
#pragma unroll
      for (int j = 0; j < 4; ++j) col[h][j] += sh.v[par][h][lane_col<T>(cg, j)];
  }
  if (ro == 0) {
#pragma unroll
    for (int h = 0; h < NRHS; ++h) {
      double* dst = colbuf + (tile * NRHS + h) * SM_B;
#pragma unroll
      for (int j = 0; j < 4; ++j) dst[lane_col<T>(cg, j)] = col[h][j];
    }
  }
}

// One tile: lane (ro, cg) of wave `wave` covers rows r0 + 4k + ro (k < ROWS,
// r0 = 64 wave / WS) and the four columns lane_col(cg, j).  T = int16_t /
// int8_t: the integers are summed against x scaled by the tile's scale sc,
// and the column sums are scaled once at the end (sc is ignored otherwise).
template <int NRHS, int WS, bool NT, typename T>
__device__ __forceinline__ void symv_tile(const T* __restrict__ tp, int64_t tile, bool diag,
                                          const double (&xr)[NRHS][SYMV_ROWS / WS],
                                          double (&rowacc)[NRHS][SYMV_ROWS / WS],
                                          const double* __restrict__ xcb, int64_t npad, int cg,
                                          int ro, int wave, int par, ColShare<NRHS, WS>& sh,
                                          double* __restrict__ colbuf, double sc = 1.0) {
  constexpr int ROWS = SYMV_ROWS / WS;
  constexpr bool QUANT = sizeof(T) <= 2;
  const int r0 = (SM_B / WS) * wave;
  double xc[NRHS][4], col[NRHS][4];
#pragma unroll
  for (int h = 0; h < NRHS; ++h)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      xc[h][j] = xcb[h * npad + lane_col<T>(cg, j)];
      if (QUANT) xc[h][j] *= sc;
      col[h][j] = 0.0;
    }
#pragma unroll
  for (int k = 0; k < ROWS; ++k) {
    const T* row = tp + (r0 + 4 * k + ro) * SM_B;
    double m[4];
    if constexpr (sizeof(T) == 8) {
      const dvec2 a = ld<NT>(reinterpret_cast<const dvec2*>(row) + cg);
      const dvec2 b = ld<NT>(reinterpret_cast<const dvec2*>(row) + 16 + cg);
      m[0] = a.x;
      m[1] = a.y;
      m[2] = b.x;
      m[3] = b.y;
    } else if constexpr (sizeof(T) == 4) {
      const fvec4 f = ld<NT>(reinterpret_cast<const fvec4*>(row) + cg);
      m[0] = f.x;
      m[1] = f.y;
      m[2] = f.z;
      m[3] = f.w;
    } else if constexpr (sizeof(T) == 2) {
      // columns 4cg .. 4cg+3: two dwords, low half first (little endian)
      const ivec2 q = ld<NT>(reinterpret_cast<const ivec2*>(row) + cg);
      m[0] = (double)(int16_t)q.x;
      m[1] = (double)(int16_t)(q.x >> 16);
      m[2] = (double)(int16_t)q.y;
      m[3] = (double)(int16_t)(q.y >> 16);
    } else {
      const int q = ld<NT>(reinterpret_cast<const int*>(row) + cg);
      m[0] = (double)(int8_t)q;
      m[1] = (double)(int8_t)(q >> 8);
      m[2] = (double)(int8_t)(q >> 16);
      m[3] = (double)(int8_t)(q >> 24);
    }
#pragma unroll
    for (int h = 0; h < NRHS; ++h) {
      rowacc[h][k] += m[0] * xc[h][0] + m[1] * xc[h][1] + m[2] * xc[h][2] + m[3] * xc[h][3];
#pragma unroll
      for (int j = 0; j < 4; ++j) col[h][j] += m[j] * xr[h][k];
    }
  }
  if (!diag) {
    if (QUANT) {
#pragma unroll
      for (int h = 0; h < NRHS; ++h)
#pragma unroll
        for (int j = 0; j < 4; ++j) col[h][j] *= sc;
    }
    store_cols<NRHS, WS, T>(col, tile, cg, ro, wave, par, sh, colbuf);
  }
}

// Optional screen of a compacted supermatrix's tiles: a tile (R, C) is
// skipped when tmax[tile] * max(xmax[R], xmax[C]) < tau, with tmax the
// tile's max |element| and xmax[B] the max |x| over the 64 packed rows of
// block B and every right-hand side -- its contribution to any element of y
// is then below 64 tau.  Zero-format tiles are skipped as well, and the
// screened reduce leaves the skipped tiles' column partials unread, so
// nothing is written for them.  tmax == nullptr: no screen (zero tiles then
// write zero column partials for the plain reduce).
struct TileScreen {
  const float* tmax;
  const float* xmax;
  float tau;
};

__device__ __forceinline__ bool tile_live(const uint8_t* __restrict__ fmt, const TileScreen& scr,
                                          int64_t tile, float xa, float xb) {
  if (fmt[tile] == TILE_ZERO) return false;
  return scr.tmax[tile] * fmaxf(xa, xb) >= scr.tau;
}

// Tiles (R, C0 .. C1-1) of one supermatrix whose tiles start at global tile
// tile0 and whose packed vector rows start at global row row0 (0 and 0 for a
// single matrix; many molecules share one launch otherwise).  Row partials
// accumulate in registers over the segment and are reduced once at the end
// by recursive halving over the 16 lanes sharing ro, into rowbuf block
// gitem.  Column partials of tile (R, C) go to colbuf block `tile` (the
// global tile index) or, GROUPED, to the wave's LDS slots: colbuf then
// points at them and the slot is C - C0.  MIXED: tiles are read through a
// format/offset directory (zero tiles skipped, FP32 / FP64 / int16 / int8,
// FP64 accumulation either way).
template <int NRHS, bool MIXED, bool NT, int WS, bool GROUPED = false>
__device__ __forceinline__ void symv_segment(const void* __restrict__ Mv,
                                             const uint8_t* __restrict__ fmt,
                                             const int64_t* __restrict__ off,
                                             const double* __restrict__ x, int64_t npad,
                                             double* __restrict__ rowbuf, double* colbuf,
                                             int64_t R, int64_t C0, int64_t C1, int64_t gitem,
                                             int64_t tile0, int64_t row0, const TileScreen& scr) {
  constexpr int ROWS = SYMV_ROWS / WS;
  const int lane = threadIdx.x & 63, cg = lane & 15, ro = lane >> 4;
  const int wave = WS == 1 ? 0 : (int)(threadIdx.x >> 6);
  const int r0 = (SM_B / WS) * wave;
  __shared__ ColShare<NRHS, WS> sh;
  int par = 0;

  double xr[NRHS][ROWS], rowacc[NRHS][ROWS];
#pragma unroll
  for (int h = 0; h < NRHS; ++h)
#pragma unroll
    for (int k = 0; k < ROWS; ++k) {
      xr[h][k] = x[h * npad + (row0 + R) * SM_B + r0 + 4 * k + ro];
      rowacc[h][k] = 0.0;
    }
  const float xmR = (MIXED && scr.tmax) ? scr.xmax[row0 + R] : 0.0f;
  // The directory entries (format, pool offset, screen) of the segment's
  // tiles are loaded at once, lane c holding tile C0 + c, and read back with
  // uniform lane reads: the tile loop then has no dependent directory load in
  // front of each tile's data loads (seg <= 64, checked on the host).
  int dir_f = TILE_ZERO;
  int64_t dir_o = 0;
  if (MIXED && lane < C1 - C0) {
    const int64_t t = tile0 + R * (R + 1) / 2 + C0 + lane;
    dir_f = fmt[t];
    if (scr.tmax && dir_f != TILE_ZERO && !tile_live(fmt, scr, t, xmR, scr.xmax[row0 + C0 + lane]))
      dir_f = TILE_ZERO + 8;  // screened out: no data, no column partials
    dir_o = off[t];
  }

  for (int64_t C = C0; C < C1; ++C) {
    const int64_t tile = tile0 + R * (R + 1) / 2 + C;
    const int64_t dtile = GROUPED ? C - C0 : tile;  // column-partial block
    const bool diag = C == R;
    const double* xcb = x + (row0 + C) * SM_B;
    if (!MIXED) {
      const double* tp = reinterpret_cast<const double*>(Mv) + (tile << (2 * SM_LOGB));
      symv_tile<NRHS, WS, NT, double>(tp, dtile, diag, xr, rowacc, xcb, npad, cg, ro, wave, par,
                                      sh, colbuf);
    } else {
      const int c = (int)(C - C0);
      const int f = __builtin_amdgcn_readlane(dir_f, c);
      if (f == TILE_ZERO + 8 || (scr.tmax && f == TILE_ZERO)) continue;
      const uint32_t olo = __builtin_amdgcn_readlane((uint32_t)dir_o, c);
      const uint32_t ohi = __builtin_amdgcn_readlane((uint32_t)(dir_o >> 32), c);
      const char* base = reinterpret_cast<const char*>(Mv) + (int64_t)(((uint64_t)ohi << 32) | olo);
      if (f == TILE_F64) {
        symv_tile<NRHS, WS, NT, double>(reinterpret_cast<const double*>(base), dtile, diag, xr,
                                        rowacc, xcb, npad, cg, ro, wave, par, sh, colbuf);
      } else if (f == TILE_F32) {
        symv_tile<NRHS, WS, NT, float>(reinterpret_cast<const float*>(base), dtile, diag, xr,
                                       rowacc, xcb, npad, cg, ro, wave, par, sh, colbuf);
      } else if (f == TILE_I16) {
        const double sc = *reinterpret_cast<const double*>(base);  // uniform: one scalar load
        symv_tile<NRHS, WS, NT, int16_t>(reinterpret_cast<const int16_t*>(base + TILE_QHDR), dtile,
                                         diag, xr, rowacc, xcb, npad, cg, ro, wave, par, sh,
                                         colbuf, sc);
      } else if (f == TILE_I8) {
        const double sc = *reinterpret_cast<const double*>(base);
        symv_tile<NRHS, WS, NT, int8_t>(reinterpret_cast<const int8_t*>(base + TILE_QHDR), dtile,
                                        diag, xr, rowacc, xcb, npad, cg, ro, wave, par, sh, colbuf,
                                        sc);
      } else if (!diag && ro == 0 && wave == 0) {
#pragma unroll
        for (int h = 0; h < NRHS; ++h) {
          dvec2 z;
          z.x = z.y = 0.0;
          dvec2* dst = reinterpret_cast<dvec2*>(colbuf + (dtile * NRHS + h) * SM_B + 4 * cg);
          dst[0] = z;
          dst[1] = z;
        }
      }
    }
    par ^= 1;
  }
  // recursive halving over the 16 lanes sharing ro (log2 ROWS steps, then
  // plain sums over the remaining lane bits): lane cg ends with the total of
  // row r0 + 4 (cg / (16 / ROWS)) + ro
  constexpr int LOGR = ROWS == 16 ? 4 : 3;
  static_assert(ROWS == 16 || ROWS == 8, "rows per lane");
#pragma unroll
  for (int h = 0; h < NRHS; ++h) {
#pragma unroll
    for (int step = 0; step < LOGR; ++step) {
      const int m = 8 >> step, half = ROWS >> (step + 1);
      const bool up = (cg & m) != 0;
#pragma unroll
      for (int i = 0; i < half; ++i) {
        const double send = up ? rowacc[h][i] : rowacc[h][i + half];
        const double keep = up ? rowacc[h][i + half] : rowacc[h][i];
        rowacc[h][i] = keep + __shfl_xor(send, m);
      }
    }
#pragma unroll
    for (int step = LOGR; step < 4; ++step) rowacc[h][0] += __shfl_xor(rowacc[h][0], 8 >> step);
    if ((cg & ((16 / ROWS) - 1)) == 0)
      rowbuf[(gitem * NRHS + h) * SM_B + r0 + 4 * (cg / (16 / ROWS)) + ro] = rowacc[h][0];
  }
}

// One wave (WS = 1) or two (WS = 2, each half of every tile's rows) per
// work item: tile row R, a segment of its tiles (see item_off).
template <int NRHS, bool MIXED, bool NT, int WS>
__device__ __forceinline__ void symv_item(const void* __restrict__ Mv, int64_t T,
                                          const uint8_t* __restrict__ fmt,
                                          const int64_t* __restrict__ off,
                                          const double* __restrict__ x, int64_t npad,
                                          double* __restrict__ rowbuf, double* __restrict__ colbuf,
                                          int seg, int64_t item, int64_t gitem, int64_t tile0,
                                          int64_t row0, const TileScreen& scr) {
  int64_t lo = 0, hi = T - 1;
  while (lo < hi) {
    int64_t mid = (lo + hi + 1) >> 1;
    if (item_off(mid, seg) <= item)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int64_t R = lo, s = item - item_off(R, seg);
  const int64_t C0 = s * seg, C1 = min(C0 + seg, R + 1);
  symv_segment<NRHS, MIXED, NT, WS>(Mv, fmt, off, x, npad, rowbuf, colbuf, R, C0, C1, gitem, tile0,
                                    row0, scr);
}

template <int NRHS, bool MIXED, bool NT, int WS>
__global__ __launch_bounds__(64 * WS, SYMV_MIN_WAVES) void symv_tiles_kernel(
    const void* __restrict__ Mv, int64_t T, const uint8_t* __restrict__ fmt,
    const int64_t* __restrict__ off, const double* __restrict__ x, int64_t npad,
    double* __restrict__ rowbuf, double* __restrict__ colbuf, int seg, TileScreen scr) {
  const int64_t item = blockIdx.x;
  symv_item<NRHS, MIXED, NT, WS>(Mv, T, fmt, off, x, npad, rowbuf, colbuf, seg, item, item, 0, 0,
                                 scr);
}

// Grouped tile pass for one right-hand side: a work-group of SYMV_G = 4
// waves takes tile rows R = 4g + w (wave w) over one column segment
// [seg s, seg s + seg) (seg a multiple of 4, at most SYMV_GSEG).  Each wave
// writes its tiles' column partials into its own LDS slots (zeros where it
// has none), and after one barrier the work-group sums the four slots of
// each column in a fixed order and writes one 512-byte block per (row group
// g, column C) instead of one per tile: the tile pass writes and the reduce
// reads a quarter of the column partials (0.83 GB per pass over the
// 1,628,110 tiles of a 480-function supermatrix).  Block (g, C) is the sum
// over the group's rows R' > C (tile (C, C) is a diagonal tile, whose column
// half the row sums already hold, and tiles (R', C) with R' < C do not
// exist); every block of the group's columns 0 .. min(4g+3, T-1) is
// written, zeros included, so the reduce needs no screen.  Group g has
// g / (seg/4) + 1 segments (its last row's), so item_off(g, seg/4) counts
// the work-groups before it.
constexpr int SYMV_G = 4;
constexpr int SYMV_GSEG = 16;

__host__ __device__ inline int64_t gcol_off(int64_t g) {  // first column block of group g
  return 2 * g * (g + 1);                                 // 4g' + 4 for each group g' < g
}
__host__ __device__ inline int64_t symv_group_blocks(int64_t T) {
  const int64_t NG = (T + SYMV_G - 1) / SYMV_G;
  return gcol_off(NG - 1) + T;  // the last group's columns 0 .. T-1
}
static bool symv_grouped(int nrhs, int seg, int ws) {
  return nrhs == 1 && ws == 1 && seg % SYMV_G == 0 && seg <= SYMV_GSEG;
}
// Row and column partials of the grouped layout, in 64-double blocks.
static void symv_group_layout(int64_t T, int seg, int64_t* items, int64_t* blocks) {
  const int64_t NG = (T + SYMV_G - 1) / SYMV_G;
  *items = item_off(NG, seg / SYMV_G);
  *blocks = *items * SYMV_G + symv_group_blocks(T);
}

template <bool MIXED, bool NT>
__global__ __launch_bounds__(64 * SYMV_G) void symv_group_kernel(
    const void* __restrict__ Mv, int64_t T, const uint8_t* __restrict__ fmt,
    const int64_t* __restrict__ off, const double* __restrict__ x, int64_t npad,
    double* __restrict__ rowbuf, double* __restrict__ colbuf, int seg, TileScreen scr) {
  __shared__ double slot[SYMV_G][SYMV_GSEG][SM_B];
  const int64_t item = blockIdx.x;
  const int sg = seg / SYMV_G;
  const int64_t NG = (T + SYMV_G - 1) / SYMV_G;
  int64_t lo = 0, hi = NG - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (item_off(mid, sg) <= item)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int64_t g = lo, C0 = (item - item_off(g, sg)) * seg;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t R = g * SYMV_G + w;
  double* mine = &slot[w][0][0];
#pragma unroll
  for (int k = 0; k < SYMV_GSEG; ++k) mine[k * SM_B + lane] = 0.0;
  if (R < T)  // C0 <= R: every row of the group reaches every segment of it
    symv_segment<1, MIXED, NT, 1, true>(Mv, fmt, off, x, npad, rowbuf, mine, R, C0,
                                        min(C0 + seg, R + 1), item * SYMV_G + w, 0, 0, scr);
  __syncthreads();
  const int64_t Rmax = min(g * SYMV_G + SYMV_G - 1, T - 1);
  const int ncol = (int)min((int64_t)seg, Rmax + 1 - C0);
  double* dst = colbuf + (gcol_off(g) + C0) * SM_B;
  for (int e = threadIdx.x; e < ncol * SM_B; e += 64 * SYMV_G) {
    const int c = e >> SM_LOGB, r = e & (SM_B - 1);
    dst[e] = ((slot[0][c][r] + slot[1][c][r]) + slot[2][c][r]) + slot[3][c][r];
  }
}

// Last index i in [0, n) with pre[i] <= v (pre ascending, pre[0] = 0).
__device__ __forceinline__ int find_seg(const int64_t* __restrict__ pre, int n, int64_t v) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= v)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Many supermatrices (one per molecule) in one persistent launch: waves
// stride over the concatenated work items of all molecules.  Per molecule m:
// T[m] tile rows, first global tile tile0[m], first global item item0[m]
// (item0[nmol] = total), first global packed row row0[m].
template <int NRHS, bool MIXED, bool NT>
__global__ __launch_bounds__(64, SYMV_MIN_WAVES) void symv_batch_kernel(
    const void* __restrict__ Mv, const uint8_t* __restrict__ fmt, const int64_t* __restrict__ off,
    const double* __restrict__ x, int64_t npad, double* __restrict__ rowbuf,
    double* __restrict__ colbuf, int seg, int nmol, const int64_t* __restrict__ molT,
    const int64_t* __restrict__ tile0, const int64_t* __restrict__ item0,
    const int64_t* __restrict__ row0) {
  const int64_t total = item0[nmol];
  const int64_t stride = (int64_t)gridDim.x * gridDim.y;
  const TileScreen none = {nullptr, nullptr, 0.0f};
  for (int64_t g = flat_block(); g < total; g += stride) {
    const int m = find_seg(item0, nmol, g);
    symv_item<NRHS, MIXED, NT, 1>(Mv, molT[m], fmt, off, x, npad, rowbuf, colbuf, seg,
                                  g - item0[m], g, tile0[m], row0[m], none);
  }
}

// Per-tile max |element| (one wave per tile).
__global__ __launch_bounds__(64) void tile_maxabs_kernel(const double* __restrict__ M,
                                                         double* __restrict__ tmax, int64_t ntiles) {
  const int64_t tile = flat_block();
  if (tile >= ntiles) return;
  const double* tp = M + (tile << (2 * SM_LOGB));
  double m = 0.0;
  for (int i = threadIdx.x; i < SM_B * SM_B; i += 64) m = fmax(m, fabs(tp[i]));
  for (int w = 32; w > 0; w >>= 1) m = fmax(m, __shfl_xor(m, w));
  if (threadIdx.x == 0) tmax[tile] = m;
}

// Copy each tile into the compacted pool in its format (int16 / int8: the
// tile's max |element| first, then the scale into the header and the
// rounded integers after it).  A lane takes eight pairs of doubles, loaded
// at once and held in registers for the max and the conversion (one read
// of the tile; the loop form with 8-byte loads and a second pass for the
// integer tiles took 15.3 ms per (H2O)_16 matrix).
__global__ __launch_bounds__(256) void tile_pack_kernel(const double* __restrict__ M,
                                                        const uint8_t* __restrict__ fmt,
                                                        const int64_t* __restrict__ off,
                                                        char* __restrict__ pool, int64_t ntiles) {
  const int64_t tile = flat_block();
  if (tile >= ntiles) return;
  const uint8_t f = fmt[tile];
  if (f == TILE_ZERO) return;
  const dvec2* tp = reinterpret_cast<const dvec2*>(M + (tile << (2 * SM_LOGB))) + threadIdx.x;
  dvec2 v[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) v[u] = __builtin_nontemporal_load(tp + 256 * u);
  char* dst = pool + off[tile];
  const int e0 = threadIdx.x;  // pair index of v[u]: e0 + 256 u (elements 2e, 2e + 1)
  if (f == TILE_F64) {
#pragma unroll
    for (int u = 0; u < 8; ++u) reinterpret_cast<dvec2*>(dst)[e0 + 256 * u] = v[u];
    return;
  }
  if (f == TILE_F32) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      fvec2 w;
      w.x = (float)v[u].x;
      w.y = (float)v[u].y;
      reinterpret_cast<fvec2*>(dst)[e0 + 256 * u] = w;
    }
    return;
  }
  __shared__ double wmax[4];
  double m = 0.0;
#pragma unroll
  for (int u = 0; u < 8; ++u) m = fmax(m, fmax(fabs(v[u].x), fabs(v[u].y)));
  for (int w = 32; w > 0; w >>= 1) m = fmax(m, __shfl_xor(m, w));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
  __syncthreads();
  m = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
  const double qmax = f == TILE_I16 ? 32767.0 : 127.0;
  // |x| * (qmax / m) <= qmax (1 + 2^-52): rint stays within the format's range
  const double inv = m > 0.0 ? qmax / m : 0.0;
  if (threadIdx.x == 0) {
    reinterpret_cast<double*>(dst)[0] = m / qmax;
    reinterpret_cast<double*>(dst)[1] = 0.0;
  }
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int q0 = (int)rint(v[u].x * inv), q1 = (int)rint(v[u].y * inv);
    if (f == TILE_I16)
      reinterpret_cast<uint32_t*>(dst + TILE_QHDR)[e0 + 256 * u] =
          (uint32_t)(uint16_t)(int16_t)q0 | ((uint32_t)(uint16_t)(int16_t)q1 << 16);
    else
      reinterpret_cast<uint16_t*>(dst + TILE_QHDR)[e0 + 256 * u] =
          (uint16_t)((uint16_t)(uint8_t)(int8_t)q0 | ((uint16_t)(uint8_t)(int8_t)q1 << 8));
  }
}

// y[h][R*64 + r] = sum of the row partials of row R's items plus the column
// partials of tiles (R', R), R' > R.  RW waves split the terms, each keeping
// eight independent sums so that eight partial-block loads are in flight per
// wave (the column partials of one row are 512-byte blocks far apart in the
// workspace, and row 0 has T - 1 of them: the reduce is latency-bound on its
// first rows).  A screened tile's block is read whether or not its pass
// wrote it (every tile has a slot) and dropped by a select, so the load does
// not wait for the screen's own loads.
constexpr int RW = 16;

__global__ __launch_bounds__(64 * RW) void symv_reduce_kernel(int nrhs, int64_t T, int64_t npad,
                                                              const double* __restrict__ rowbuf,
                                                              const double* __restrict__ colbuf,
                                                              double* __restrict__ y, int seg,
                                                              const uint8_t* __restrict__ fmt,
                                                              TileScreen scr) {
  const int64_t R = blockIdx.x;
  const int h = blockIdx.y;
  const int r = threadIdx.x & 63, w = threadIdx.x >> 6;
  double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int64_t i0 = item_off(R, seg), i1 = item_off(R + 1, seg);
  for (int64_t it = i0 + w; it < i1; it += RW) a[0] += rowbuf[(it * nrhs + h) * SM_B + r];
  const float xmR = scr.tmax ? scr.xmax[R] : 0.0f;
  // the tile screen (tile_live, taken at a column block), without short-circuits:
  // every load issued
  auto col = [&](int64_t Rp, auto screened) -> double {
    const int64_t tile = Rp * (Rp + 1) / 2 + R;
    const double v = __builtin_nontemporal_load(colbuf + (tile * nrhs + h) * SM_B + r);
    if constexpr (decltype(screened)::value) {
      const bool live = (fmt[tile] != TILE_ZERO) & (scr.tmax[tile] * fmaxf(scr.xmax[Rp], xmR) >= scr.tau);
      return live ? v : 0.0;
    } else {
      return v;
    }
  };
  auto sweep = [&](auto screened) {
    int64_t Rp = R + 1 + w;
    for (; Rp + 7 * RW < T; Rp += 8 * RW) {
      // eight blocks' loads (partials, formats, maxima) first, then the sums
      double v[8];
      int f[8];
      float tm[8], xm[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int64_t Ru = Rp + u * RW, tile = Ru * (Ru + 1) / 2 + R;
        v[u] = __builtin_nontemporal_load(colbuf + (tile * nrhs + h) * SM_B + r);
        if constexpr (decltype(screened)::value) {
          f[u] = fmt[tile];
          tm[u] = scr.tmax[tile];
          xm[u] = scr.xmax[Ru];
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if constexpr (decltype(screened)::value) {
          const bool live = (f[u] != TILE_ZERO) & (tm[u] * fmaxf(xm[u], xmR) >= scr.tau);
          a[u] += live ? v[u] : 0.0;
        } else {
          a[u] += v[u];
        }
      }
    }
    for (; Rp < T; Rp += RW) a[0] += col(Rp, screened);
  };
  if (scr.tmax)
    sweep(std::true_type{});
  else
    sweep(std::false_type{});
  __shared__ double part[RW][64];
  part[w][r] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  __syncthreads();
  if (w == 0) {
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < RW; ++k) v += part[k][r];
    y[h * npad + R * SM_B + r] = v;
  }
}

// Reduce of the grouped layout: y[R] = the row partials of row R's
// work-groups plus the column blocks (g', R) of the groups g' >= R / 4.
__global__ __launch_bounds__(64 * RW) void symv_group_reduce_kernel(int64_t T, int64_t npad,
                                                                    const double* __restrict__ rowbuf,
                                                                    const double* __restrict__ colbuf,
                                                                    double* __restrict__ y, int seg) {
  const int64_t R = blockIdx.x;
  const int r = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int sg = seg / SYMV_G;
  const int64_t g = R / SYMV_G, NG = (T + SYMV_G - 1) / SYMV_G;
  const int wr = (int)(R - g * SYMV_G);
  double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int64_t i0 = item_off(g, sg), i1 = item_off(g + 1, sg);
  for (int64_t it = i0 + w; it < i1; it += RW) a[0] += rowbuf[(it * SYMV_G + wr) * SM_B + r];
  int64_t gp = g + w;
  for (; gp + 7 * RW < NG; gp += 8 * RW) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      v[u] = __builtin_nontemporal_load(colbuf + (gcol_off(gp + u * RW) + R) * SM_B + r);
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] += v[u];
  }
  for (; gp < NG; gp += RW) a[1] += colbuf[(gcol_off(gp) + R) * SM_B + r];
  __shared__ double part[RW][64];
  part[w][r] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  __syncthreads();
  if (w == 0) {
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < RW; ++k) v += part[k][r];
    y[R * SM_B + r] = v;
  }
}

// Batched reduce: one block per global packed row g (all molecules).
__global__ __launch_bounds__(256) void symv_batch_reduce_kernel(
    int nrhs, int64_t npad, const double* __restrict__ rowbuf, const double* __restrict__ colbuf,
    double* __restrict__ y, int seg, int nmol, const int64_t* __restrict__ molT,
    const int64_t* __restrict__ tile0, const int64_t* __restrict__ item0,
    const int64_t* __restrict__ row0) {
  const int64_t g = blockIdx.x;
  const int h = blockIdx.y;
  const int r = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int m = find_seg(row0, nmol, g);
  const int64_t R = g - row0[m], T = molT[m];
  double acc = 0.0;
  const int64_t i0 = item0[m] + item_off(R, seg), i1 = item0[m] + item_off(R + 1, seg);
  for (int64_t it = i0 + w; it < i1; it += 4) acc += rowbuf[(it * nrhs + h) * SM_B + r];
  for (int64_t Rp = R + 1 + w; Rp < T; Rp += 4) {
    const int64_t tile = tile0[m] + Rp * (Rp + 1) / 2 + R;
    acc += colbuf[(tile * nrhs + h) * SM_B + r];
  }
  __shared__ double part[4][64];
  part[w][r] = acc;
  __syncthreads();
  if (w == 0) y[h * npad + g * SM_B + r] = part[0][r] + part[1][r] + part[2][r] + part[3][r];
}

// D: [nrhs][nmol][nmax][nmax] (molecule m uses its leading nbf[m] block) ->
// packed x: [nrhs][npad], molecule m at rows row0[m] .. row0[m+1].
__global__ void batch_pack_density_kernel(const double* __restrict__ D, int64_t nmax, int nmol,
                                          const int64_t* __restrict__ nbf,
                                          const int64_t* __restrict__ row0,
                                          double* __restrict__ x, int64_t npad) {
  const int64_t pg = flat_block() * blockDim.x + threadIdx.x;
  const int h = blockIdx.z;
  if (pg >= npad) return;
  const int m = find_seg(row0, nmol, pg / SM_B);
  const int64_t p = pg - row0[m] * SM_B, n = nbf[m], npair = n * (n + 1) / 2;
  double v = 0.0;
  if (p < npair) {
    const int64_t k = tri_root(p), l = p - k * (k + 1) / 2;
    v = D[(((int64_t)h * nmol + m) * nmax + k) * nmax + l] * (k == l ? 1.0 : 2.0);
  }
  x[h * npad + pg] = v;
}

// F[h][m][i][j] = H[m][i][j] + scale * y[h][row0[m]*64 + pair(i, j)] inside
// molecule m's nbf[m] block, H[m][i][j] outside it.
__global__ void batch_unpack_fock_kernel(const double* __restrict__ H, const double* __restrict__ y,
                                         double scale, int64_t nmax, int nmol,
                                         const int64_t* __restrict__ nbf,
                                         const int64_t* __restrict__ row0, int64_t npad,
                                         double* __restrict__ F) {
  const int64_t e = flat_block() * blockDim.x + threadIdx.x;
  const int h = blockIdx.z;
  const int64_t per = nmax * nmax;
  if (e >= nmol * per) return;
  const int64_t m = e / per, ij = e - m * per, i = ij / nmax, j = ij - i * nmax;
  double v = H ? H[e] : 0.0;
  if (i < nbf[m] && j < nbf[m]) v += scale * y[h * npad + row0[m] * SM_B + pidx(i, j)];
  F[(int64_t)h * nmol * per + e] = v;
}

__global__ void pack_density_kernel(const double* __restrict__ D, int64_t nbf, int nrhs,
                                    double* __restrict__ x, int64_t npad) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int h = blockIdx.y;
  if (p >= npad) return;
  const int64_t npair = nbf * (nbf + 1) / 2;
  double v = 0.0;
  if (p < npair) {
    const int64_t k = tri_root(p), l = p - k * (k + 1) / 2;
    v = D[(h * nbf + k) * nbf + l] * (k == l ? 1.0 : 2.0);
  }
  x[h * npad + p] = v;
}

// xmax[B] = max over the right-hand sides and the 64 rows of block B of |x|.
__global__ __launch_bounds__(64) void block_maxabs_kernel(const double* __restrict__ x, int nrhs,
                                                         int64_t npad, float* __restrict__ xmax) {
  const int64_t B = blockIdx.x;
  double m = 0.0;
  for (int h = 0; h < nrhs; ++h) m = fmax(m, fabs(x[h * npad + B * SM_B + threadIdx.x]));
  for (int w = 32; w > 0; w >>= 1) m = fmax(m, __shfl_xor(m, w));
  if (threadIdx.x == 0) xmax[B] = (float)m;
}

__global__ void unpack_fock_kernel(const double* __restrict__ H, const double* __restrict__ y,
                                   double scale, int64_t nbf, int64_t npad,
                                   double* __restrict__ F) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int h = blockIdx.y;
  if (e >= nbf * nbf) return;
  const int64_t i = e / nbf, j = e % nbf;
  double v = scale * y[h * npad + pidx(i, j)];
  if (H) v += H[e];
  F[h * nbf * nbf + e] = v;
}

// Dense J, K: one block per (m, n), reduction over (l, s).
__global__ __launch_bounds__(256) void dense_jk_kernel(const double* __restrict__ eri,
                                                       const double* __restrict__ D, int64_t n,
                                                       double* __restrict__ J,
                                                       double* __restrict__ K) {
  const int64_t m = blockIdx.x / n, nn = blockIdx.x % n;
  double sj = 0.0, sk = 0.0;
  for (int64_t ls = threadIdx.x; ls < n * n; ls += blockDim.x) {
    const int64_t l = ls / n, s = ls % n;
    const double d = D[ls];
    sj += d * eri[((m * n + nn) * n + l) * n + s];
    sk += d * eri[((m * n + s) * n + l) * n + nn];
  }
  __shared__ double bj[256], bk[256];
  bj[threadIdx.x] = sj;
  bk[threadIdx.x] = sk;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      bj[threadIdx.x] += bj[threadIdx.x + w];
      bk[threadIdx.x] += bk[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    J[m * n + nn] = bj[0];
    K[m * n + nn] = bk[0];
  }
}

// Called by qc_init with `device` current.
int symv_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices)
    return set_error("qc_init: device %d beyond the %d the scratch table holds", device, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return 0;
  void* p = nullptr;
  if (check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")) return 1;
  g_scratch_ptr[device] = p;
  return 0;
}

}  // namespace qcmd

using namespace qcmd;

extern "C" {

int qc_sm_assemble_maxabs(int kind, const double* Jsm, double* out, int64_t nbf, double* tmax,
                          void* stream) {
  if (kind != 0 && kind != 1) return set_error("qc_sm_assemble: kind %d", kind);
  if (nbf <= 0) return set_error("qc_sm_assemble: nbf %lld", (long long)nbf);
  if (!Jsm || !out) return set_error("qc_sm_assemble: null supermatrix");
  if (Jsm == out) return set_error("qc_sm_assemble: in-place assembly is not supported");
  const int64_t T = sm_tiles_per_dim(nbf);
  const int64_t ntiles = T * (T + 1) / 2;
  hipLaunchKernelGGL(sm_assemble_kernel, grid_1d(ntiles, 256), dim3(256), 0, (hipStream_t)stream,
                     kind, Jsm, out, nbf, ntiles, tmax);
  return check(hipGetLastError(), "qc_sm_assemble launch");
}

int qc_sm_assemble(int kind, const double* Jsm, double* out, int64_t nbf, void* stream) {
  return qc_sm_assemble_maxabs(kind, Jsm, out, nbf, nullptr, stream);
}

int qc_sm_p_from_jk(double* J, const double* K, int64_t nbf, double* tmax, void* stream) {
  if (nbf <= 0) return set_error("qc_sm_p_from_jk: nbf %lld", (long long)nbf);
  if (!J || !K || J == K) return set_error("qc_sm_p_from_jk: J and K must be distinct buffers");
  const int64_t T = sm_tiles_per_dim(nbf);
  const int64_t ntiles = T * (T + 1) / 2;
  hipLaunchKernelGGL(sm_p_from_jk_kernel, grid_1d(ntiles, 256), dim3(256), 0, (hipStream_t)stream,
                     J, K, ntiles, tmax);
  return check(hipGetLastError(), "qc_sm_p_from_jk launch");
}

int qc_symv_config(int seg, int nontemporal) {
  // a work item's directory is held one tile per lane
  if (seg < 1 || seg > 64) return set_error("qc_symv_config: seg %d outside [1, 64]", seg);
  g_symv_seg = seg;
  g_symv_nt = nontemporal ? 1 : 0;
  return 0;
}

int qc_symv_split(int nrhs, int waves) {
  if (nrhs < 1 || nrhs > 4) return set_error("qc_symv_split: nrhs %d outside [1, 4]", nrhs);
  if (waves != 1 && waves != 2) return set_error("qc_symv_split: waves %d (1 or 2)", waves);
  g_symv_ws[nrhs] = waves;
  return 0;
}

// Workspace of the layout a call with this configuration uses.
static int64_t symv_need_bytes(int64_t nbf, int nrhs) {
  if (symv_grouped(nrhs, g_symv_seg, g_symv_ws[1])) {
    int64_t items, blocks;
    symv_group_layout(sm_tiles_per_dim(nbf), g_symv_seg, &items, &blocks);
    return blocks * SM_B * (int64_t)sizeof(double);
  }
  return symv_work_bytes(nbf, nrhs, g_symv_seg);
}

int64_t qc_symv_workspace_bytes(int64_t nbf, int nrhs) {
  // sized for any later qc_symv_config / qc_symv_split: the per-tile layout
  // at segment 1 and, for one right-hand side, the grouped layout at its
  // smallest segment (4), which needs more for a single tile row (4 row
  // blocks per work-group)
  int64_t need = symv_work_bytes(nbf, nrhs, 1);
  if (nrhs == 1) {
    int64_t items, blocks;
    symv_group_layout(sm_tiles_per_dim(nbf), SYMV_G, &items, &blocks);
    const int64_t g = blocks * SM_B * (int64_t)sizeof(double);
    if (g > need) need = g;
  }
  return need;
}

static int symv_phase(int phase, const double* M, int64_t nbf, int nrhs, const double* x,
                      double* y, void* work, void* stream, const uint8_t* fmt = nullptr,
                      const int64_t* off = nullptr,
                      TileScreen scr = TileScreen{nullptr, nullptr, 0.0f}) {
  if (nrhs < 1 || nrhs > 4) return set_error("qc_symv: nrhs %d outside [1, 4]", nrhs);
  if (nbf <= 0) return set_error("qc_symv: nbf %lld", (long long)nbf);
  if (!M || !x || !work || (phase != 1 && !y)) return set_error("qc_symv: null pointer");
  const int64_t T = sm_tiles_per_dim(nbf), npad = T * SM_B;
  const int seg = g_symv_seg;
  const bool grouped = symv_grouped(nrhs, seg, g_symv_ws[1]);
  int64_t items = item_off(T, seg), gblocks = 0;
  if (grouped) symv_group_layout(T, seg, &items, &gblocks);
  double* rowbuf = (double*)work;
  double* colbuf = rowbuf + items * (grouped ? SYMV_G : nrhs) * SM_B;
  hipStream_t s = (hipStream_t)stream;
  if ((phase & 1) && grouped) {
    if (fmt) {
      if (g_symv_nt)
        hipLaunchKernelGGL((symv_group_kernel<true, true>), dim3((unsigned)items), dim3(64 * SYMV_G),
                           0, s, (const void*)M, T, fmt, off, x, npad, rowbuf, colbuf, seg, scr);
      else
        hipLaunchKernelGGL((symv_group_kernel<true, false>), dim3((unsigned)items),
                           dim3(64 * SYMV_G), 0, s, (const void*)M, T, fmt, off, x, npad, rowbuf,
                           colbuf, seg, scr);
    } else {
      if (g_symv_nt)
        hipLaunchKernelGGL((symv_group_kernel<false, true>), dim3((unsigned)items),
                           dim3(64 * SYMV_G), 0, s, (const void*)M, T, fmt, off, x, npad, rowbuf,
                           colbuf, seg, scr);
      else
        hipLaunchKernelGGL((symv_group_kernel<false, false>), dim3((unsigned)items),
                           dim3(64 * SYMV_G), 0, s, (const void*)M, T, fmt, off, x, npad, rowbuf,
                           colbuf, seg, scr);
    }
    if (check(hipGetLastError(), "qc_symv tiles")) return 1;
  } else if (phase & 1) {
    switch (nrhs) {
#define QC_L(N, MX, NT)                                                                        \
  do {                                                                                         \
    if (g_symv_ws[N] == 2)                                                                     \
      hipLaunchKernelGGL((symv_tiles_kernel<N, MX, NT, 2>), dim3((unsigned)items), dim3(128), 0, \
                         s, (const void*)M, T, fmt, off, x, npad, rowbuf, colbuf, seg, scr);    \
    else                                                                                       \
      hipLaunchKernelGGL((symv_tiles_kernel<N, MX, NT, 1>), dim3((unsigned)items), dim3(64), 0,  \
                         s, (const void*)M, T, fmt, off, x, npad, rowbuf, colbuf, seg, scr);    \
  } while (0)
#define QC_S(N)                     \
  case N:                           \
    if (fmt) {                      \
      if (g_symv_nt)                \
        QC_L(N, true, true);        \
      else                          \
        QC_L(N, true, false);       \
    } else {                        \
      if (g_symv_nt)                \
        QC_L(N, false, true);       \
      else                          \
        QC_L(N, false, false);      \
    }                               \
    break;
      QC_S(1) QC_S(2) QC_S(3) QC_S(4)
#undef QC_S
#undef QC_L
    }
    if (check(hipGetLastError(), "qc_symv tiles")) return 1;
  }
  if ((phase & 2) && grouped) {
    hipLaunchKernelGGL(symv_group_reduce_kernel, dim3((unsigned)T), dim3(64 * RW), 0, s, T, npad,
                       rowbuf, colbuf, y, seg);
    if (check(hipGetLastError(), "qc_symv reduce")) return 1;
  } else if (phase & 2) {
    hipLaunchKernelGGL(symv_reduce_kernel, dim3((unsigned)T, nrhs), dim3(64 * RW), 0, s, nrhs, T,
                       npad, rowbuf, colbuf, y, seg, fmt, scr);
    if (check(hipGetLastError(), "qc_symv reduce")) return 1;
  }
  return 0;
}

int qc_symv(const double* M, int64_t nbf, int nrhs, const double* x, double* y, void* work,
            void* stream) {
  return symv_phase(3, M, nbf, nrhs, x, y, work, stream);
}

int qc_symv_noalloc(const double* M, int64_t nbf, int nrhs, const double* x, double* y,
                    void* stream) {
  if (nrhs < 1 || nrhs > 4) return set_error("qc_symv_noalloc: nrhs %d outside [1, 4]", nrhs);
  if (nbf <= 0) return set_error("qc_symv_noalloc: nbf %lld", (long long)nbf);
  const int64_t need = symv_need_bytes(nbf, nrhs);
  if (need > kSymvScratchBytes)
    return set_error("qc_symv_noalloc: nbf %lld with %d right-hand sides needs %lld B of "
                     "workspace, above the %lld B scratch (use qc_symv with a caller workspace)",
                     (long long)nbf, nrhs, (long long)need, (long long)kSymvScratchBytes);
  int dev = -1;
  if (check(hipGetDevice(&dev), "qc_symv_noalloc: hipGetDevice")) return 1;
  void* w = symv_scratch_ptr(dev);
  if (!w) return set_error("qc_symv_noalloc: qc_init not called on device %d", dev);
  return symv_phase(3, M, nbf, nrhs, x, y, w, stream);
}

int qc_symv_tiles(const double* M, int64_t nbf, int nrhs, const double* x, void* work,
                  void* stream) {
  return symv_phase(1, M, nbf, nrhs, x, nullptr, work, stream);
}

int qc_symv_mixed_tiles(const void* pool, const uint8_t* fmt, const int64_t* off, int64_t nbf,
                        int nrhs, const double* x, void* work, void* stream) {
  if (!fmt || !off) return set_error("qc_symv_mixed_tiles: null tile directory");
  return symv_phase(1, (const double*)pool, nbf, nrhs, x, nullptr, work, stream, fmt, off);
}

int qc_block_maxabs(const double* x, int64_t nbf, int nrhs, float* xmax, void* stream) {
  if (nrhs < 1 || nrhs > 4) return set_error("qc_block_maxabs: nrhs %d outside [1, 4]", nrhs);
  if (nbf <= 0 || !x || !xmax) return set_error("qc_block_maxabs: bad arguments");
  const int64_t T = sm_tiles_per_dim(nbf), npad = T * SM_B;
  hipLaunchKernelGGL(block_maxabs_kernel, dim3((unsigned)T), dim3(64), 0, (hipStream_t)stream, x,
                     nrhs, npad, xmax);
  return check(hipGetLastError(), "qc_block_maxabs");
}

static int screen_args(const char* who, const uint8_t* fmt, const float* tmax,
                       const float* xmax, double tau) {
  if (!fmt || !tmax || !xmax) return set_error("%s: null tile directory or maxima", who);
  if (!(tau >= 0.0)) return set_error("%s: tau %g must be >= 0", who, tau);
  return 0;
}

int qc_symv_screened_tiles(const void* pool, const uint8_t* fmt, const int64_t* off,
                           const float* tmax, int64_t nbf, int nrhs, const double* x,
                           const float* xmax, double tau, void* work, void* stream) {
  if (!off) return set_error("qc_symv_screened_tiles: null tile offsets");
  if (screen_args("qc_symv_screened_tiles", fmt, tmax, xmax, tau)) return 1;
  return symv_phase(1, (const double*)pool, nbf, nrhs, x, nullptr, work, stream, fmt, off,
                    TileScreen{tmax, xmax, (float)tau});
}

int qc_symv_screened_reduce(int64_t nbf, int nrhs, double* y, void* work, const uint8_t* fmt,
                            const float* tmax, const float* xmax, double tau, void* stream) {
  if (screen_args("qc_symv_screened_reduce", fmt, tmax, xmax, tau)) return 1;
  return symv_phase(2, (const double*)work, nbf, nrhs, (const double*)work, y, work, stream, fmt,
                    nullptr, TileScreen{tmax, xmax, (float)tau});
}

int qc_tiles_maxabs(const double* M, int64_t ntiles, double* tmax, void* stream) {
  if (ntiles <= 0) return 0;
  if (!M || !tmax) return set_error("qc_tiles_maxabs: null pointer");
  hipLaunchKernelGGL(tile_maxabs_kernel, grid_1d(ntiles, 64), dim3(64), 0, (hipStream_t)stream, M,
                     tmax, ntiles);
  return check(hipGetLastError(), "qc_tiles_maxabs");
}

int qc_tiles_pack(const double* M, int64_t ntiles, const uint8_t* fmt, const int64_t* off,
                  void* pool, void* stream) {
  if (ntiles <= 0) return 0;
  if (!M || !fmt || !off || !pool) return set_error("qc_tiles_pack: null pointer");
  hipLaunchKernelGGL(tile_pack_kernel, grid_1d(ntiles, 256), dim3(256), 0, (hipStream_t)stream, M,
                     fmt, off, (char*)pool, ntiles);
  return check(hipGetLastError(), "qc_tiles_pack");
}

int64_t qc_batch_items(const int64_t* molT, int nmol, int seg, int64_t* item0) {
  if (!molT || !item0 || nmol < 0 || seg < 1) return -1;
  int64_t acc = 0;
  for (int m = 0; m < nmol; ++m) {
    item0[m] = acc;
    acc += item_off(molT[m], seg);
  }
  item0[nmol] = acc;
  return acc;
}

int qc_batch_symv_tiles(const void* M, const uint8_t* fmt, const int64_t* off, int nmol,
                        const int64_t* molT, const int64_t* tile0, const int64_t* item0,
                        const int64_t* row0, int64_t nitems, int64_t ntiles, int64_t npad,
                        int nrhs, int seg, const double* x, void* work, void* stream) {
  (void)ntiles;
  if (nrhs < 1 || nrhs > 4) return set_error("qc_batch_symv_tiles: nrhs %d outside [1, 4]", nrhs);
  if (seg < 1 || seg > 64) return set_error("qc_batch_symv_tiles: seg %d outside [1, 64]", seg);
  if (!M || !molT || !tile0 || !item0 || !row0 || !x || !work || (fmt && !off))
    return set_error("qc_batch_symv_tiles: null pointer");
  if (nmol <= 0 || nitems <= 0) return 0;
  double* rowbuf = (double*)work;
  double* colbuf = rowbuf + nitems * nrhs * SM_B;
  // persistent: enough waves to fill the chip several times over, each
  // striding over the items of every molecule
  const int64_t waves = nitems < 16384 ? nitems : 16384;
  hipStream_t s = (hipStream_t)stream;
  switch (nrhs) {
#define QC_L(N, MX, NT)                                                                           \
  hipLaunchKernelGGL((symv_batch_kernel<N, MX, NT>), grid_1d(waves, 64), dim3(64), 0, s, M, fmt, \
                     off, x, npad, rowbuf, colbuf, seg, nmol, molT, tile0, item0, row0)
#define QC_S(N)              \
  case N:                    \
    if (fmt) {               \
      if (g_symv_nt)         \
        QC_L(N, true, true);  \
      else                   \
        QC_L(N, true, false); \
    } else {                 \
      if (g_symv_nt)         \
        QC_L(N, false, true); \
      else                   \
        QC_L(N, false, false); \
    }                        \
    break;
    QC_S(1) QC_S(2) QC_S(3) QC_S(4)
#undef QC_S
#undef QC_L
  }
  return check(hipGetLastError(), "qc_batch_symv_tiles");
}

int qc_batch_symv_reduce(int nmol, const int64_t* molT, const int64_t* tile0, const int64_t* item0,
                         const int64_t* row0, int64_t nrows, int64_t nitems, int64_t npad, int nrhs,
                         int seg, double* y, void* work, void* stream) {
  if (nrhs < 1 || nrhs > 4) return set_error("qc_batch_symv_reduce: nrhs %d outside [1, 4]", nrhs);
  if (!molT || !tile0 || !item0 || !row0 || !y || !work)
    return set_error("qc_batch_symv_reduce: null pointer");
  if (nmol <= 0 || nrows <= 0) return 0;
  if (nrows * SM_B > npad) return set_error("qc_batch_symv_reduce: %lld rows exceed npad", (long long)nrows);
  const double* rowbuf = (const double*)work;
  const double* colbuf = rowbuf + nitems * nrhs * SM_B;
  hipLaunchKernelGGL(symv_batch_reduce_kernel, dim3((unsigned)nrows, nrhs), dim3(256), 0,
                     (hipStream_t)stream, nrhs, npad, rowbuf, colbuf, y, seg, nmol, molT, tile0,
                     item0, row0);
  return check(hipGetLastError(), "qc_batch_symv_reduce");
}

int qc_batch_pack_density(const double* D, int64_t nmax, int nmol, const int64_t* nbf,
                          const int64_t* row0, int nrhs, double* x, int64_t npad, void* stream) {
  if (!D || !nbf || !row0 || !x) return set_error("qc_batch_pack_density: null pointer");
  if (nrhs < 1 || nrhs > 4 || nmol <= 0) return set_error("qc_batch_pack_density: nrhs %d nmol %d", nrhs, nmol);
  dim3 g = grid_1d((npad + 255) / 256, 256);
  g.z = nrhs;
  hipLaunchKernelGGL(batch_pack_density_kernel, g, dim3(256), 0, (hipStream_t)stream, D, nmax, nmol,
                     nbf, row0, x, npad);
  return check(hipGetLastError(), "qc_batch_pack_density");
}

int qc_batch_unpack_fock(const double* H, const double* y, double scale, int64_t nmax, int nmol,
                         const int64_t* nbf, const int64_t* row0, int nrhs, int64_t npad, double* F,
                         void* stream) {
  if (!y || !nbf || !row0 || !F) return set_error("qc_batch_unpack_fock: null pointer");
  if (nrhs < 1 || nrhs > 4 || nmol <= 0) return set_error("qc_batch_unpack_fock: nrhs %d nmol %d", nrhs, nmol);
  dim3 g = grid_1d((nmol * nmax * nmax + 255) / 256, 256);
  g.z = nrhs;
  hipLaunchKernelGGL(batch_unpack_fock_kernel, g, dim3(256), 0, (hipStream_t)stream, H, y, scale,
                     nmax, nmol, nbf, row0, npad, F);
  return check(hipGetLastError(), "qc_batch_unpack_fock");
}

int qc_sm_tile_maxabs(const double* M, int64_t nbf, double* tmax, void* stream) {
  const int64_t T = sm_tiles_per_dim(nbf);
  const int64_t nt = T * (T + 1) / 2;
  hipLaunchKernelGGL(tile_maxabs_kernel, grid_1d(nt, 64), dim3(64), 0, (hipStream_t)stream, M,
                     tmax, nt);
  return check(hipGetLastError(), "qc_sm_tile_maxabs");
}

int qc_sm_pack(const double* M, int64_t nbf, const uint8_t* fmt, const int64_t* off, void* pool,
               void* stream) {
  const int64_t T = sm_tiles_per_dim(nbf);
  const int64_t nt = T * (T + 1) / 2;
  hipLaunchKernelGGL(tile_pack_kernel, grid_1d(nt, 256), dim3(256), 0, (hipStream_t)stream, M,
                     fmt, off, (char*)pool, nt);
  return check(hipGetLastError(), "qc_sm_pack");
}

int qc_symv_reduce(int64_t nbf, int nrhs, double* y, void* work, void* stream) {
  // the reduction reads only the workspace; pass work as a dummy M/x
  return symv_phase(2, (const double*)work, nbf, nrhs, (const double*)work, y, work, stream);
}

int qc_pack_density(const double* D, int64_t nbf, int nrhs, double* x, void* stream) {
  if (nrhs < 1) return set_error("qc_pack_density: nrhs %d", nrhs);
  const int64_t npad = sm_tiles_per_dim(nbf) * SM_B;
  hipLaunchKernelGGL(pack_density_kernel, dim3((unsigned)((npad + 255) / 256), nrhs), dim3(256), 0,
                     (hipStream_t)stream, D, nbf, nrhs, x, npad);
  return check(hipGetLastError(), "qc_pack_density");
}

int qc_unpack_fock(const double* H, const double* y, double scale, int64_t nbf, int nrhs,
                   double* F, void* stream) {
  if (nrhs < 1) return set_error("qc_unpack_fock: nrhs %d", nrhs);
  const int64_t npad = sm_tiles_per_dim(nbf) * SM_B;
  hipLaunchKernelGGL(unpack_fock_kernel, dim3((unsigned)((nbf * nbf + 255) / 256), nrhs),
                     dim3(256), 0, (hipStream_t)stream, H, y, scale, nbf, npad, F);
  return check(hipGetLastError(), "qc_unpack_fock");
}

int qc_dense_jk(const double* eri, const double* D, int64_t nbf, double* J, double* K,
                void* stream) {
  if (nbf <= 0 || nbf > 4096) return set_error("qc_dense_jk: nbf %lld", (long long)nbf);
  hipLaunchKernelGGL(dense_jk_kernel, dim3((unsigned)(nbf * nbf)), dim3(256), 0,
                     (hipStream_t)stream, eri, D, nbf, J, K);
  return check(hipGetLastError(), "qc_dense_jk");
}

}  // extern "C"
