// Instantiation of the engine kernels for storage mode M_F16.
#include "kernels.hpp"

namespace nlos {
template struct Launch<M_F16>;
}  // namespace nlos
