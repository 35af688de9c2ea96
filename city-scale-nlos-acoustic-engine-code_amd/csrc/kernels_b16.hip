// Instantiation of the engine kernels for storage mode M_B16.
#include "kernels.hpp"

namespace nlos {
template struct Launch<M_B16>;
}  // namespace nlos
