// Kernel instances for bra class 5 (generated list, one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_bra5(ClassKernels (&tab)[6][6]) {
    reg_class<5, 0>(tab);
    reg_class<5, 1>(tab);
    reg_class<5, 2>(tab);
    reg_class<5, 3>(tab);
}
}  // namespace tlno_amd
