// Team-evaluator instances for bra class 12 (a pair with a g shell; one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_hi12a(HighKernels (&tab)[kNumClasses - kNumLowClasses][kNumClasses]) {
    reg_high<12, 0>(tab);
    reg_high<12, 1>(tab);
    reg_high<12, 2>(tab);
    reg_high<12, 3>(tab);
    reg_high<12, 4>(tab);
    reg_high<12, 5>(tab);
}
}  // namespace tlno_amd
