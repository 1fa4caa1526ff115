// Team-evaluator instances for bra class 7 (a pair with an f shell; one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_hi7(HighKernels (&tab)[kNumClasses - kNumLowClasses][kNumClasses]) {
    reg_high<7, 0>(tab);
    reg_high<7, 1>(tab);
    reg_high<7, 2>(tab);
    reg_high<7, 3>(tab);
    reg_high<7, 4>(tab);
    reg_high<7, 5>(tab);
    reg_high<7, 6>(tab);
    reg_high<7, 7>(tab);
}
}  // namespace tlno_amd
