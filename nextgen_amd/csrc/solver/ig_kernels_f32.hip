// Mixed-precision instantiation of the ideal-gas kernels (R = float), compiled
// with -ffp-contract=on: a product and a sum within one expression fuse into an
// FMA, the same way in every kernel variant (contraction across statements,
// -ffp-contract=fast, fuses differently in differently shaped variants, and
// results would then depend on the variant chosen); the FP64 parts keep
// contraction off inside their own bodies.
#include "ig_kernels.inl"

namespace cgk {
CGK_INSTANTIATE(float)
}  // namespace cgk
