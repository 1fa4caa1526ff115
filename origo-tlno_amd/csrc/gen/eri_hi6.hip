// Team-evaluator instances for bra class 6 (a pair with an f shell; one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_hi6(HighKernels (&tab)[kNumClasses - kNumLowClasses][kNumClasses]) {
    reg_high<6, 0>(tab);
    reg_high<6, 1>(tab);
    reg_high<6, 2>(tab);
    reg_high<6, 3>(tab);
    reg_high<6, 4>(tab);
    reg_high<6, 5>(tab);
    reg_high<6, 6>(tab);
}
}  // namespace tlno_amd
