// Team-evaluator instances for bra class 9 (a pair with an f shell; one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_hi9a(HighKernels (&tab)[kNumClasses - kNumLowClasses][kNumClasses]) {
    reg_high<9, 0>(tab);
    reg_high<9, 1>(tab);
    reg_high<9, 2>(tab);
    reg_high<9, 3>(tab);
    reg_high<9, 4>(tab);
}
}  // namespace tlno_amd
