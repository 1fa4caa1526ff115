// Team-evaluator instances for bra class 10 (a pair with a g shell; one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_hi10(HighKernels (&tab)[kNumClasses - kNumLowClasses][kNumClasses]) {
    reg_high<10, 0>(tab);
    reg_high<10, 1>(tab);
    reg_high<10, 2>(tab);
    reg_high<10, 3>(tab);
    reg_high<10, 4>(tab);
    reg_high<10, 5>(tab);
    reg_high<10, 6>(tab);
    reg_high<10, 7>(tab);
    reg_high<10, 8>(tab);
    reg_high<10, 9>(tab);
    reg_high<10, 10>(tab);
}
}  // namespace tlno_amd
