// Team-evaluator instances for bra class 8 (a pair with an f shell; one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_hi8a(HighKernels (&tab)[kNumClasses - kNumLowClasses][kNumClasses]) {
    reg_high<8, 0>(tab);
    reg_high<8, 1>(tab);
    reg_high<8, 2>(tab);
    reg_high<8, 3>(tab);
    reg_high<8, 4>(tab);
}
}  // namespace tlno_amd
