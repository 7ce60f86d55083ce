// instantiation of the y-direction kernels for chunk size M = 12
#include "loma_yops_impl.h"
namespace loma {
LOMA_INSTANTIATE_M(12)
}  // namespace loma
