// Kernel instances for bra class 4 (generated list, one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_bra4b(ClassKernels (&tab)[6][6]) {
    reg_class<4, 4>(tab);
}
}  // namespace tlno_amd
