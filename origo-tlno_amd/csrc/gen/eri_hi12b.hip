// Team-evaluator instances for bra class 12 (a pair with a g shell; one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_hi12b(HighKernels (&tab)[kNumClasses - kNumLowClasses][kNumClasses]) {
    reg_high<12, 6>(tab);
    reg_high<12, 7>(tab);
    reg_high<12, 8>(tab);
    reg_high<12, 9>(tab);
}
}  // namespace tlno_amd
