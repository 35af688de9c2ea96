// Instantiation of the engine kernels for storage mode M_F64.
#include "kernels.hpp"

namespace nlos {
template struct Launch<M_F64>;
}  // namespace nlos
