// Team-evaluator instances for bra class 12 (a pair with a g shell; one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_hi12c(HighKernels (&tab)[kNumClasses - kNumLowClasses][kNumClasses]) {
    reg_high<12, 10>(tab);
    reg_high<12, 11>(tab);
    reg_high<12, 12>(tab);
}
}  // namespace tlno_amd
