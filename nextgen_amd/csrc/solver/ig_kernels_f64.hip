// FP64 instantiation of the ideal-gas kernels (and the precision-independent
// launchers), compiled with -ffp-contract=off.
#define CGK_COMMON
#include "ig_kernels.inl"

namespace cgk {
CGK_INSTANTIATE(double)
}  // namespace cgk
