// FP32-form instantiation of the field kernels (storage float, arithmetic
// FP64, one rounding per stored value). Built with contraction allowed:
// unlike the plain form, nothing here is compared bitwise with the
// reference.
#include "fields_impl.hpp"

namespace aprl_gpu {

APRL_FIELDS_INSTANTIATE(float)

}  // namespace aprl_gpu
