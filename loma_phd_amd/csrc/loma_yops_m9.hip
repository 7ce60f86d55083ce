// instantiation of the y-direction kernels for chunk size M = 9
#include "loma_yops_impl.h"
namespace loma {
LOMA_INSTANTIATE_M(9)
}  // namespace loma
