// Kernel instances for bra class 2 (generated list, one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_bra2(ClassKernels (&tab)[6][6]) {
    reg_class<2, 0>(tab);
    reg_class<2, 1>(tab);
    reg_class<2, 2>(tab);
}
}  // namespace tlno_amd
