// qcmd_tables.h -- compile-time component tables (indexable with unrolled
// loop counters, so every lookup folds to a constant).
#pragma once

namespace qcmd {

// Cartesian exponents of component k of shell l (l <= 3), ordered lx
// descending, then ly descending (jax_qc/integrals/mcmurchie_davidson.py:39-50).
__device__ constexpr int CX[4][10] = {{0}, {1, 0, 0}, {2, 1, 1, 0, 0, 0}, {3, 2, 2, 1, 1, 1, 0, 0, 0, 0}};
__device__ constexpr int CY[4][10] = {{0}, {0, 1, 0}, {0, 1, 0, 2, 1, 0}, {0, 1, 0, 2, 1, 0, 3, 2, 1, 0}};
__device__ constexpr int CZ[4][10] = {{0}, {0, 0, 1}, {0, 0, 1, 0, 1, 2}, {0, 0, 1, 0, 1, 2, 0, 1, 2, 3}};

// Cartesian -> spherical rows for l <= 2 in libcint order
// (jax_qc/integrals/spherical.py:34-67).  [l][row][cart]
#define QC_R3 1.7320508075688772935
__device__ constexpr double SPH[3][5][6] = {
    {{1.0}},
    {{1.0, 0, 0}, {0, 1.0, 0}, {0, 0, 1.0}},
    {{0, QC_R3, 0, 0, 0, 0},
     {0, 0, 0, 0, QC_R3, 0},
     {-0.5, 0, 0, -0.5, 0, 1.0},
     {0, 0, QC_R3, 0, 0, 0},
     {0.5 * QC_R3, 0, 0, -0.5 * QC_R3, 0, 0}}};
#undef QC_R3

}  // namespace qcmd
