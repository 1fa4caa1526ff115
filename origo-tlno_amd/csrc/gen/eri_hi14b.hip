// Team-evaluator instances for bra class 14 (a pair with a g shell; one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_hi14b(HighKernels (&tab)[kNumClasses - kNumLowClasses][kNumClasses]) {
    reg_high<14, 6>(tab);
    reg_high<14, 7>(tab);
    reg_high<14, 8>(tab);
    reg_high<14, 9>(tab);
}
}  // namespace tlno_amd
