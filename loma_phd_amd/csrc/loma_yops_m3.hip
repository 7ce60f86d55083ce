// instantiation of the y-direction kernels for chunk size M = 3
#include "loma_yops_impl.h"
namespace loma {
LOMA_INSTANTIATE_M(3)
}  // namespace loma
