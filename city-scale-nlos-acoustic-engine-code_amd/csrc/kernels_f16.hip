// Instantiation of the engine kernels for storage mode M_F16, and the measurement kernel of
// the FP16 two-array sweep's access pattern.
#include "kernels.hpp"

namespace nlos {
template struct Launch<M_F16>;

// The data movement of the FP16 two-array sweep and nothing else: every interior cell loads
// slot q of x - e_q from `in` and stores it, non-temporally, into slot q of x in `out`
// (blocked arrays, 256x1 row segments, the sweep's lane addressing).  No mask, bounce-back,
// collision, probes or records: the time is what the access pattern itself allows.
__global__ __launch_bounds__(256) void k_pull_pattern16(const _Float16 *__restrict__ in, _Float16 *__restrict__ out,
                                                        GridDesc g, int tiles_x) {
    const int b = blockIdx.x;
    const int bi = b % tiles_x;
    const int row = b / tiles_x;
    const int j = row % g.ny + 1;
    const int k = row / g.ny + 1;
    const int i0 = bi * 256 + 1;
    if (i0 + (int)threadIdx.x > g.nx) return;
    const long long cb = cell_index(g, i0, j, k);
    const Lanes<M_F16> L(cb, threadIdx.x, g.ncell);
    _Float16 v[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) v[q] = glb(in)[L.from(g, q, q)];
#pragma unroll
    for (int q = 0; q < Q; ++q) __builtin_nontemporal_store(v[q], glb(out) + L.self(q));
}

hipError_t launch_pull_pattern16(const void *in, void *out, const GridDesc &g, hipStream_t s) {
    const int tiles_x = (g.nx + 255) / 256;
    const long long blocks = (long long)tiles_x * g.ny * g.nz;
    k_pull_pattern16<<<(unsigned)blocks, 256, 0, s>>>(static_cast<const _Float16 *>(in), static_cast<_Float16 *>(out),
                                                       g, tiles_x);
    return hipGetLastError();
}
}  // namespace nlos
