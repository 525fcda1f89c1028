// qcmd_internal.h -- host-side helpers shared by the translation units.
#pragma once

#include <hip/hip_runtime.h>

namespace qcmd {

int set_error(const char* fmt, ...);
int check(hipError_t e, const char* what);
const double* boys_table64();
const float* boys_table32();
int symv_scratch_init(int device);  // jk.hip: device address of the qc_symv_noalloc scratch

// The 21 ERI classes (la lb | lc ld), la >= lb, lc >= ld, pair class
// la(la+1)/2+lb >= lc(lc+1)/2+ld, l <= 2.
#define QC_ERI_CLASSES(X)                                                            \
  X(0, 0, 0, 0) X(1, 0, 0, 0) X(1, 0, 1, 0) X(1, 1, 0, 0) X(1, 1, 1, 0) X(1, 1, 1, 1) \
  X(2, 0, 0, 0) X(2, 0, 1, 0) X(2, 0, 1, 1) X(2, 0, 2, 0) X(2, 1, 0, 0) X(2, 1, 1, 0) \
  X(2, 1, 1, 1) X(2, 1, 2, 0) X(2, 1, 2, 1) X(2, 2, 0, 0) X(2, 2, 1, 0) X(2, 2, 1, 1) \
  X(2, 2, 2, 0) X(2, 2, 2, 1) X(2, 2, 2, 2)

}  // namespace qcmd
