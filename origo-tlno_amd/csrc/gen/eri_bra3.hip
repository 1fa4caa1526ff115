// Kernel instances for bra class 3 (generated list, one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_bra3(ClassKernels (&tab)[6][6]) {
    reg_class<3, 0>(tab);
    reg_class<3, 1>(tab);
    reg_class<3, 2>(tab);
    reg_class<3, 3>(tab);
}
}  // namespace tlno_amd
