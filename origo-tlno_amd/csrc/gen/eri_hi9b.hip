// Team-evaluator instances for bra class 9 (a pair with an f shell; one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_hi9b(HighKernels (&tab)[kNumClasses - kNumLowClasses][kNumClasses]) {
    reg_high<9, 5>(tab);
    reg_high<9, 6>(tab);
    reg_high<9, 7>(tab);
}
}  // namespace tlno_amd
