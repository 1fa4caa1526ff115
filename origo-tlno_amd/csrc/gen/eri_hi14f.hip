// Team-evaluator instances for bra class 14 (a pair with a g shell; one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_hi14f(HighKernels (&tab)[kNumClasses - kNumLowClasses][kNumClasses]) {
    reg_high<14, 14>(tab);
}
}  // namespace tlno_amd
