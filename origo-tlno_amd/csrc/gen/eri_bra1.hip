// Kernel instances for bra class 1 (generated list, one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_bra1(ClassKernels (&tab)[6][6]) {
    reg_class<1, 0>(tab);
    reg_class<1, 1>(tab);
}
}  // namespace tlno_amd
