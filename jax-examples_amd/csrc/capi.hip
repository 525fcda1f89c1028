// capi.hip -- C-ABI entry points of libqcmd.so: per-device setup (Boys
// tables), error reporting, the Boys kernel and the ERI class dispatch.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>

#include <mutex>
#include <vector>

#include "qcmd_device.h"
#include "qcmd_internal.h"
#include "qcmd_types.h"

namespace qcmd {

static thread_local char g_err[512] = "";

int set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return 1;
}

int check(hipError_t e, const char* what) {
  if (e == hipSuccess) return 0;
  return set_error("%s: %s", what, hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// Boys tables.  F_n(T_i) for T_i = i h, n < BOYS_NCOL, from the series
//   F_n(T) = exp(-T) sum_k (2T)^k / ((2n+1)(2n+3)...(2n+2k+1))
// (all terms positive) in long double for the top column, then the downward
// recursion F_n = (2T F_{n+1} + exp(-T)) / (2n+1), also in long double.
// ---------------------------------------------------------------------------
static void boys_host_table(std::vector<double>& t64, std::vector<float>& t32) {
  t64.assign((size_t)BOYS_NGRID * BOYS_NCOL, 0.0);
  t32.assign((size_t)BOYS_NGRID * BOYS_NCOL, 0.0f);
  for (int i = 0; i < BOYS_NGRID; ++i) {
    long double T = (long double)i * (long double)BOYS_H;
    int ntop = BOYS_NCOL - 1;
    long double term = 1.0L / (2 * ntop + 1), sum = term;
    for (int k = 1; k < 4000; ++k) {
      term *= 2.0L * T / (long double)(2 * ntop + 2 * k + 1);
      sum += term;
      if (term < sum * 1e-22L) break;
    }
    long double e = expl(-T);
    long double F[BOYS_NCOL];
    F[ntop] = e * sum;
    for (int n = ntop - 1; n >= 0; --n) F[n] = (2.0L * T * F[n + 1] + e) / (2 * n + 1);
    for (int n = 0; n < BOYS_NCOL; ++n) {
      t64[(size_t)i * BOYS_NCOL + n] = (double)F[n];
      t32[(size_t)i * BOYS_NCOL + n] = (float)F[n];
    }
  }
}

struct DeviceState {
  double* boys64 = nullptr;
  float* boys32 = nullptr;
};
static std::mutex g_mu;
static std::vector<DeviceState> g_dev;

const DeviceState* device_state() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  if (dev < 0 || dev >= (int)g_dev.size() || !g_dev[dev].boys64) return nullptr;
  return &g_dev[dev];
}

const double* boys_table64() {
  const DeviceState* s = device_state();
  return s ? s->boys64 : nullptr;
}
const float* boys_table32() {
  const DeviceState* s = device_state();
  return s ? s->boys32 : nullptr;
}

template <int N, typename Real>
__global__ void boys_kernel(const Real* __restrict__ t, Real* __restrict__ F, int64_t n,
                            const Real* __restrict__ tab) {
  int64_t i = flat_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Real f[N + 1];
  boys_fn<N, Real>(t[i], tab, f);
#pragma unroll
  for (int k = 0; k <= N; ++k) F[i * (N + 1) + k] = f[k];
}

template <typename Real>
static int launch_boys(int nmax, const Real* t, Real* F, int64_t n, hipStream_t s,
                       const Real* tab) {
  if (n <= 0) return 0;
  dim3 grid = grid_1d((n + 255) / 256, 256), block(256);
  switch (nmax) {
#define QC_B(N) \
  case N: hipLaunchKernelGGL((boys_kernel<N, Real>), grid, block, 0, s, t, F, n, tab); break;
    QC_B(0) QC_B(1) QC_B(2) QC_B(3) QC_B(4) QC_B(5) QC_B(6) QC_B(7) QC_B(8) QC_B(9) QC_B(10)
    QC_B(11) QC_B(12)
#undef QC_B
    default:
      return set_error("qc_boys: nmax=%d outside [0, %d]", nmax, BOYS_NMAX);
  }
  return check(hipGetLastError(), "qc_boys launch");
}

// ERI class launchers, one per translation unit (generated by build.py).
#define QC_DECL(A, B, C, D)                                                   \
  hipError_t eri_launch_##A##B##C##D##_f64(const EriArgs& a, hipStream_t s);  \
  hipError_t eri_launch_##A##B##C##D##_f32(const EriArgs& a, hipStream_t s);  \
  hipError_t eri_launchc_##A##B##C##D##_f64(const EriArgs& a, hipStream_t s); \
  hipError_t eri_launchc_##A##B##C##D##_f32(const EriArgs& a, hipStream_t s);
QC_ERI_CLASSES(QC_DECL)
#undef QC_DECL

}  // namespace qcmd

using namespace qcmd;

extern "C" {

const char* qc_last_error(void) { return g_err; }

int qc_init(int device) {
  int ndev = 0;
  if (check(hipGetDeviceCount(&ndev), "hipGetDeviceCount")) return 1;
  if (device < 0 || device >= ndev) return set_error("qc_init: device %d of %d", device, ndev);
  if (check(hipSetDevice(device), "hipSetDevice")) return 1;
  std::lock_guard<std::mutex> lk(g_mu);
  if ((int)g_dev.size() < ndev) g_dev.resize(ndev);
  DeviceState& st = g_dev[device];
  if (symv_scratch_init(device)) return 1;
  if (st.boys64) return 0;
  std::vector<double> t64;
  std::vector<float> t32;
  boys_host_table(t64, t32);
  if (check(hipMalloc(&st.boys64, t64.size() * sizeof(double)), "hipMalloc boys64")) return 1;
  if (check(hipMalloc(&st.boys32, t32.size() * sizeof(float)), "hipMalloc boys32")) return 1;
  if (check(hipMemcpy(st.boys64, t64.data(), t64.size() * sizeof(double), hipMemcpyHostToDevice),
            "boys64 upload"))
    return 1;
  if (check(hipMemcpy(st.boys32, t32.data(), t32.size() * sizeof(float), hipMemcpyHostToDevice),
            "boys32 upload"))
    return 1;
  return 0;
}

int qc_boys(int nmax, const double* t, double* F, int64_t n, void* stream) {
  const double* tab = boys_table64();
  if (!tab) return set_error("qc_boys: qc_init not called on the current device");
  return launch_boys<double>(nmax, t, F, n, (hipStream_t)stream, tab);
}

int qc_boys_f32(int nmax, const float* t, float* F, int64_t n, void* stream) {
  const float* tab = boys_table32();
  if (!tab) return set_error("qc_boys_f32: qc_init not called on the current device");
  return launch_boys<float>(nmax, t, F, n, (hipStream_t)stream, tab);
}

int qc_eri(int cab, int ccd, int fp32, const qc_eri_args* args_in, void* stream) {
  if (!args_in) return set_error("qc_eri: null args");
  if (cab < 0 || cab > 5 || ccd < 0 || ccd > cab)
    return set_error("qc_eri: pair classes (%d, %d) invalid (need 0 <= ccd <= cab <= 5)", cab, ccd);
  qc_eri_args a = *args_in;
  a.boys64 = boys_table64();
  a.boys32 = boys_table32();
  if (!a.boys64) return set_error("qc_eri: qc_init not called on the current device");
  if (fp32 && !a.pairs32) return set_error("qc_eri: fp32 launch without an FP32 pair table");
  if (a.mode < 0 || a.mode > 2) return set_error("qc_eri: mode %d", a.mode);
  if (a.mode == ERI_SCHWARZ && cab != ccd) return set_error("qc_eri: Schwarz mode needs cab == ccd");
  if (a.nquartets < 0 || (a.nquartets > 0 && a.nbra <= 0))
    return set_error("qc_eri: nquartets %lld with nbra %d", (long long)a.nquartets, a.nbra);
  if (a.variant < 0 || a.variant > 2) return set_error("qc_eri: variant %d", a.variant);
  if (a.block != 0 && a.block != 64 && a.block != 128)
    return set_error("qc_eri: block %d (need 0, 64 or 128)", a.block);
  if (a.sp_base && a.mode != ERI_SUPER) return set_error("qc_eri: sp_base needs mode 1");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipErrorInvalidValue;
  // variant 0: the form measured faster on MI355X ((H2O)_8 even-tempered, FP64):
  // one wave per quartet for (dp|dp), (dd|pp), (dd|dp), (dd|dd), one thread
  // per quartet for every other class (e.g. (dp|pp) 14.9 vs 78 ms)
  const bool wave_default = (cab == 4 && ccd == 4) || (cab == 5 && (ccd == 2 || ccd == 4 || ccd == 5));
#define QC_CALL(A, B, C, D)                                                                    \
  if (cab == (A) * ((A) + 1) / 2 + (B) && ccd == (C) * ((C) + 1) / 2 + (D)) {                  \
    const bool coop = a.variant == 2 || (a.variant == 0 && wave_default);                      \
    if (coop)                                                                                  \
      e = fp32 ? eri_launchc_##A##B##C##D##_f32(a, s) : eri_launchc_##A##B##C##D##_f64(a, s); \
    else                                                                                       \
      e = fp32 ? eri_launch_##A##B##C##D##_f32(a, s) : eri_launch_##A##B##C##D##_f64(a, s);   \
  }
  QC_ERI_CLASSES(QC_CALL)
#undef QC_CALL
  return check(e, "qc_eri launch");
}

}  // extern "C"
