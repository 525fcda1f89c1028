// qcmd_types.h -- argument blocks shared by the host launchers and kernels.
#pragma once

#include <stdint.h>

#include "qcmd.h"

namespace qcmd {

// One primitive pair in the pair table (AoS, 12 reals, 96 B in FP64):
//   p, Px, Py, Pz, PAx, PAy, PAz, PBx, PBy, PBz, K, alpha
// K = c_a c_b exp(-alpha beta / p |A-B|^2) (contraction coefficients folded
// in); alpha is the exponent on the first (larger-l) centre.
constexpr int PP_STRIDE = 12;
enum { PP_P = 0, PP_PX = 1, PP_PAX = 4, PP_PBX = 7, PP_K = 10, PP_ALPHA = 11 };

// Tile-major packed supermatrix: basis-function pairs p = i(i+1)/2 + j
// (i >= j) index rows and columns; 64 x 64 tiles (R, C) with C <= R stored
// contiguously in the order R(R+1)/2 + C, each tile row-major.  Diagonal
// tiles hold both triangles.
constexpr int SM_LOGB = 6;
constexpr int SM_B = 1 << SM_LOGB;

enum EriMode : int { ERI_DENSE = 0, ERI_SUPER = 1, ERI_SCHWARZ = 2 };

using EriArgs = qc_eri_args;

}  // namespace qcmd
