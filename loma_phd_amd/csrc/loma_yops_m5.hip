// instantiation of the y-direction kernels for chunk size M = 5
#include "loma_yops_impl.h"
namespace loma {
LOMA_INSTANTIATE_M(5)
}  // namespace loma
