// Instantiation of the engine kernels for storage mode M_F32.
#include "kernels.hpp"

namespace nlos {
template struct Launch<M_F32>;
}  // namespace nlos
