// instantiation of the y-direction kernels for chunk size M = 10
#include "loma_yops_impl.h"
namespace loma {
LOMA_INSTANTIATE_M(10)
}  // namespace loma
