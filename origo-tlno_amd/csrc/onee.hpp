// One-electron S/T/V launch interface (see onee.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "eri_device.hpp"

namespace tlno_amd {

// Writes the bra-oriented Cartesian blocks of every shell pair into the
// internal-order matrices S, T, V (M x M, zeroed by the caller); blocks of
// diagonal pairs are halved so that X + X^T is the full matrix.
void launch_onee(const PairData& pd, const DevTables& tb, const int* dL, const double* dexp, const double* dR,
                 const double* dZ, int nnuc, int npair, int M, double* S, double* T, double* V, hipStream_t st);

}  // namespace tlno_amd
