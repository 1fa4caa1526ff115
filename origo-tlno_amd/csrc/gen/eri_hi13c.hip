// Team-evaluator instances for bra class 13 (a pair with a g shell; one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_hi13c(HighKernels (&tab)[kNumClasses - kNumLowClasses][kNumClasses]) {
    reg_high<13, 10>(tab);
    reg_high<13, 11>(tab);
}
}  // namespace tlno_amd
