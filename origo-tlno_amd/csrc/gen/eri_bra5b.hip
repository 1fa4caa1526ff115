// Kernel instances for bra class 5 (generated list, one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_bra5b(ClassKernels (&tab)[6][6]) {
    reg_class<5, 4>(tab);
}
}  // namespace tlno_amd
