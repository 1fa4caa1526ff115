// Team-evaluator instances for bra class 14 (a pair with a g shell; one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_hi14a(HighKernels (&tab)[kNumClasses - kNumLowClasses][kNumClasses]) {
    reg_high<14, 0>(tab);
    reg_high<14, 1>(tab);
    reg_high<14, 2>(tab);
    reg_high<14, 3>(tab);
    reg_high<14, 4>(tab);
    reg_high<14, 5>(tab);
}
}  // namespace tlno_amd
