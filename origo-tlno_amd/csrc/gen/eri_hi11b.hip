// Team-evaluator instances for bra class 11 (a pair with a g shell; one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_hi11b(HighKernels (&tab)[kNumClasses - kNumLowClasses][kNumClasses]) {
    reg_high<11, 6>(tab);
    reg_high<11, 7>(tab);
    reg_high<11, 8>(tab);
    reg_high<11, 9>(tab);
    reg_high<11, 10>(tab);
    reg_high<11, 11>(tab);
}
}  // namespace tlno_amd
