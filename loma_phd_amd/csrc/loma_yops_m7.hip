// instantiation of the y-direction kernels for chunk size M = 7
#include "loma_yops_impl.h"
namespace loma {
LOMA_INSTANTIATE_M(7)
}  // namespace loma
