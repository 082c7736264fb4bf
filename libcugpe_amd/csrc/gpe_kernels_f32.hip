// complex64 instantiation of the engine kernels
#include "gpe_kernels_impl.hpp"

namespace gpe {
GPE_INSTANTIATE(float)
}  // namespace gpe
