// Kernel instances for bra class 5 (generated list, one TU per group of classes).
#include "../eri_register.hpp"
namespace tlno_amd {
void register_kernels_bra5c(ClassKernels (&tab)[6][6]) {
    reg_class<5, 5>(tab);
}
}  // namespace tlno_amd
