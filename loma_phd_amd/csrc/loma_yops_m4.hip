// instantiation of the y-direction kernels for chunk size M = 4
#include "loma_yops_impl.h"
namespace loma {
LOMA_INSTANTIATE_M(4)
}  // namespace loma
