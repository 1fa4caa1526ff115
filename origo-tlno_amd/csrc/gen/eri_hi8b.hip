// Team-evaluator instances for bra class 8 (a pair with an f shell; one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_hi8b(HighKernels (&tab)[kNumClasses - kNumLowClasses][kNumClasses]) {
    reg_high<8, 5>(tab);
    reg_high<8, 6>(tab);
    reg_high<8, 7>(tab);
    reg_high<8, 8>(tab);
}
}  // namespace tlno_amd
