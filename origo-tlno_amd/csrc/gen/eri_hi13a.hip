// Team-evaluator instances for bra class 13 (a pair with a g shell; one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_hi13a(HighKernels (&tab)[kNumClasses - kNumLowClasses][kNumClasses]) {
    reg_high<13, 0>(tab);
    reg_high<13, 1>(tab);
    reg_high<13, 2>(tab);
    reg_high<13, 3>(tab);
    reg_high<13, 4>(tab);
    reg_high<13, 5>(tab);
}
}  // namespace tlno_amd
