// instantiation of the y-direction kernels for chunk size M = 8
#include "loma_yops_impl.h"
namespace loma {
LOMA_INSTANTIATE_M(8)
}  // namespace loma
