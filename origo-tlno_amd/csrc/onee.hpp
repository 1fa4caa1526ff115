// One-electron S/T/V launch interface (see onee.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "eri_device.hpp"

namespace tlno_amd {

// Writes the bra-oriented Cartesian blocks of the listed shell pairs into the
// internal-order matrices S, T, V (M x M, zeroed by the caller); blocks of
// diagonal pairs are halved so that X + X^T is the full matrix.  list holds
// {pair, V wanted} entries grouped by pair class (class_begin[0..15]); vthr is
// the nuclear primitive-pair screen threshold (0 = off).
void launch_onee(const PairData& pd, const DevTables& tb, const int* dL, const double* dexp, const double* dR,
                 const double* dZ, int nnuc, const int2* dlist, const int* class_begin, double vthr, int M,
                 double* S, double* T, double* V, hipStream_t st);

// max angular normalization factor over the components of a complete shell
double onee_namax(int L);

}  // namespace tlno_amd
