// Kernel instances for bra class 4 (generated list, one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_bra4(ClassKernels (&tab)[6][6]) {
    reg_class<4, 0>(tab);
    reg_class<4, 1>(tab);
    reg_class<4, 2>(tab);
    reg_class<4, 3>(tab);
}
}  // namespace tlno_amd
