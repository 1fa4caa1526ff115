// Team-evaluator instances for bra class 11 (a pair with a g shell; one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_hi11a(HighKernels (&tab)[kNumClasses - kNumLowClasses][kNumClasses]) {
    reg_high<11, 0>(tab);
    reg_high<11, 1>(tab);
    reg_high<11, 2>(tab);
    reg_high<11, 3>(tab);
    reg_high<11, 4>(tab);
    reg_high<11, 5>(tab);
}
}  // namespace tlno_amd
