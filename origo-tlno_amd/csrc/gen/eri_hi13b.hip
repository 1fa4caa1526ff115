// Team-evaluator instances for bra class 13 (a pair with a g shell; one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_hi13b(HighKernels (&tab)[kNumClasses - kNumLowClasses][kNumClasses]) {
    reg_high<13, 6>(tab);
    reg_high<13, 7>(tab);
    reg_high<13, 8>(tab);
    reg_high<13, 9>(tab);
}
}  // namespace tlno_amd
