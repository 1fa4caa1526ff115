// Boys-function entry points of the C++ facade (reference:
// core/math/general_math.hpp:6-22), evaluated by the device Boys evaluator of
// the ERI kernels.
#ifndef TLNO_AMD_CORE_MATH_GENERAL_MATH_HPP
#define TLNO_AMD_CORE_MATH_GENERAL_MATH_HPP

namespace tlno::core::math {

// F_order(x)
double boys_single(int order, double x);
// out[n] = (-2p)^n F_n(x), n = 0..max_n
void boys_r000_all(int max_n, double x, double p, double* out);
// W lanes sharing p, out[n * W + lane]; W = 1, 2, 4 or 8
template <int W>
void boys_r000_all_batch(int max_n, const double* xs, double p, double* out);
// lanes whose bit is clear in active_mask output zeros
template <int W>
void boys_r000_all_batch_masked(int max_n, const double* xs, double p, double* out, unsigned active_mask);
void boys_r000_all_x4(int max_n, const double* x4, double p, double* out);

}  // namespace tlno::core::math

#endif
