// Kernel instances for bra class 0 (generated list, one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_bra0(ClassKernels (&tab)[6][6]) {
    reg_class<0, 0>(tab);
}
}  // namespace tlno_amd
