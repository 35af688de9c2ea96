// Family-A step (01_acoustic_scattering.cpp / 04_hpc_scalability.cpp step_pull semantics)
// for the engine's legacy mode, FP64 with contraction off so that every value rounds as in
// the reference.  Four launches per step (five with 04's line source, k_a_line, first):
//   k_a_halo_xy   x and y halo cells at rest (no neighbour rank, 01:234-246)
//   k_a_halo_z    z halo planes copied periodically, x/y halo columns included (01:283-300)
//   k_a_collide   post-collision value of every fluid cell, halo included (01:315-343: the
//                 reference collides the source cell once per destination; the moments
//                 and the update do not depend on which destination asks)
//   k_a_stream    interior fluid cells: pull the post-collision value, or bounce the cell's
//                 own pre-collision value back when the source is solid, blend the sponge
//                 of the destination, rho = sum in slot order (01:308-356)
// Both passes store non-temporally (nothing re-reads the lines from cache): 5 980 -> 6 085
// MLUPS on 04's 208^3 geometry.
// Populations in ordering A (01:17-19), slot-major as every FP64 array: f[q*ncell + cell],
// cell as in GridDesc.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "launch.hpp"

namespace nlos {

namespace {

constexpr int EXA[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int EYA[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int EZA[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
constexpr int INVA[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};

// Slot of ordering B (common.hpp) with the velocity of slot q of ordering A.
constexpr int a2b(int q) {
    for (int b = 0; b < Q; ++b)
        if (EX[b] == EXA[q] && EY[b] == EYA[q] && EZ[b] == EZA[q]) return b;
    return -1;
}
constexpr int A2B[Q] = {a2b(0), a2b(1), a2b(2), a2b(3), a2b(4), a2b(5), a2b(6), a2b(7), a2b(8), a2b(9),
                        a2b(10), a2b(11), a2b(12), a2b(13), a2b(14), a2b(15), a2b(16), a2b(17), a2b(18)};
static_assert(A2B[0] == 0 && A2B[18] >= 0 && A2B[7] == 7, "ordering A to B");

// get_feq (01:86-91), including the products with zero velocity components.
__device__ __forceinline__ double feq_a(int k, double r, double ux, double uy, double uz) {
    const double cu = (double)EXA[k] * ux + (double)EYA[k] * uy + (double)EZA[k] * uz;
    const double u2 = ux * ux + uy * uy + uz * uz;
    return W64[k] * r * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * u2);
}

__device__ __forceinline__ long long off_a(const GridDesc &g, int q) {
    return (long long)EXA[q] + (long long)EYA[q] * g.px + (long long)EZA[q] * g.plane;
}

}  // namespace

__global__ __launch_bounds__(256) void k_a_halo_xy(double *f, GridDesc g) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long ring = 2LL * (g.nx + 2) + 2LL * g.ny;   // cells with i or j on the halo
    if (t >= ring * g.nz) return;
    const int k = (int)(t / ring) + 1;
    long long r = t % ring;
    int i, j;
    if (r < g.nx + 2) { i = (int)r; j = 0; }
    else if ((r -= g.nx + 2) < g.nx + 2) { i = (int)r; j = g.ny + 1; }
    else if ((r -= g.nx + 2) < g.ny) { i = 0; j = (int)r + 1; }
    else { i = g.nx + 1; j = (int)(r - g.ny) + 1; }
    const long long c = cell_index(g, i, j, k);
    // feq(q, 1, 0, 0, 0) = w_q exactly (every term of the bracket but 1 is zero)
#pragma unroll
    for (int q = 0; q < Q; ++q) f[pop_index<M_F64>(q, c, g.ncell)] = W64[q];
}

__global__ __launch_bounds__(256) void k_a_halo_z(double *f, GridDesc g) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long plane = (long long)(g.nx + 2) * (g.ny + 2);
    if (t >= plane) return;
    const int i = (int)(t % (g.nx + 2)), j = (int)(t / (g.nx + 2));
    const long long lo = cell_index(g, i, j, 0), hi = cell_index(g, i, j, g.nz + 1);
    const long long top = cell_index(g, i, j, g.nz), bot = cell_index(g, i, j, 1);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        f[pop_index<M_F64>(q, lo, g.ncell)] = f[pop_index<M_F64>(q, top, g.ncell)];
        f[pop_index<M_F64>(q, hi, g.ncell)] = f[pop_index<M_F64>(q, bot, g.ncell)];
    }
}

__global__ __launch_bounds__(256) void k_a_collide(const double *f, double *post, const uint8_t *mask, GridDesc g,
                                                   double tau) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)(g.nx + 2) * (g.ny + 2) * (g.nz + 2);
    if (t >= total) return;
    const int i = (int)(t % (g.nx + 2));
    const int j = (int)((t / (g.nx + 2)) % (g.ny + 2));
    const int k = (int)(t / ((long long)(g.nx + 2) * (g.ny + 2)));
    const long long c = cell_index(g, i, j, k);
    if (mask[c] & MK_SOLID) return;
    double v[Q];
#pragma unroll
    for (int p = 0; p < Q; ++p) v[p] = f[pop_index<M_F64>(p, c, g.ncell)];
    double sr = 0.0, sux = 0.0, suy = 0.0, suz = 0.0;
#pragma unroll
    for (int p = 0; p < Q; ++p) {
        sr += v[p];
        sux += v[p] * (double)EXA[p];
        suy += v[p] * (double)EYA[p];
        suz += v[p] * (double)EZA[p];
    }
    if (sr == 0.0) sr = 1.0;
    sux /= sr;
    suy /= sr;
    suz /= sr;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const double fe = feq_a(q, sr, sux, suy, suz);
        __builtin_nontemporal_store(v[q] - (v[q] - fe) / tau, post + pop_index<M_F64>(q, c, g.ncell));
    }
}

__global__ __launch_bounds__(256) void k_a_stream(AStepArgs a) {
    const GridDesc &g = a.g;
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)g.nx * g.ny * g.nz;
    if (n >= total) return;
    const int i = (int)(n % g.nx) + 1;
    const int j = (int)((n / g.nx) % g.ny) + 1;
    const int k = (int)(n / ((long long)g.nx * g.ny)) + 1;
    const long long c = cell_index(g, i, j, k);
    const uint8_t m = a.mask[c];
    if (m & MK_SOLID) {
        if (a.field) a.field[n] = 1.0;   // 01:367 writes rho = 1 for solid cells
        return;
    }
    const double s = a.sigma[n];
    const bool in_sponge = s > 1e-6;
    double out[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const long long src = c - off_a(g, q);
        if (a.mask[src] & MK_SOLID) {
            out[q] = a.f[pop_index<M_F64>(INVA[q], c, g.ncell)];
        } else {
            const double p = a.post[pop_index<M_F64>(q, src, g.ncell)];
            out[q] = in_sponge ? (1.0 - s) * p + s * W64[q] : p;   // feq(q, 1, 0, 0, 0) = w_q
        }
    }
    double sum = 0.0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        __builtin_nontemporal_store(out[q], a.f + pop_index<M_F64>(q, c, g.ncell));
        sum += out[q];
    }
    if (!isfinite(sum)) atomicOr(a.flag, 1);
    if (a.field) a.field[n] = sum;
    if ((m & MK_PROBE) && a.probe_row) {
        int lo = 0, hi = a.n_probe_cells - 1;
        while (lo <= hi) {
            const int mid = (lo + hi) >> 1;
            const long long v = a.probe_cells[mid];
            if (v == c) { a.probe_row[mid] = sum - 1.0; break; }
            if (v < c) lo = mid + 1; else hi = mid - 1;
        }
    }
}

// The stream of k_a_stream and, for every step but the last of a run, the next step's
// line-source add (k_a_line) and collision (k_a_collide) of the same cell, in one pass: the
// post-collision values are pulled straight from the previous pass, so the state itself is
// written once, at the end of the run (617 -> about 313 algorithmic bytes per update).  Every
// value is formed by the same operations in the same order as in the separate kernels.
__global__ __launch_bounds__(256) void k_a_fused(AFusedArgs a) {
    const GridDesc &g = a.g;
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)g.nx * g.ny * g.nz;
    if (n >= total) return;
    const int i = (int)(n % g.nx) + 1;
    const int j = (int)((n / g.nx) % g.ny) + 1;
    const int k = (int)(n / ((long long)g.nx * g.ny)) + 1;
    const long long c = cell_index(g, i, j, k);
    const uint8_t m = a.mask[c];
    if (m & MK_SOLID) {
        if (a.field) a.field[n] = 1.0;   // 01:367 writes rho = 1 for solid cells
        return;
    }
    const double s = a.sigma[n];
    const bool in_sponge = s > 1e-6;
    // bit b of the wall bits: c - e_b (ordering B) is solid; only cells next to a wall read them
    const uint32_t wb = (m & MK_NEAR) ? a.wall_bits[c] : 0u;
    const bool wall = wb != 0u;
    double out[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const long long src = c - off_a(g, q);
        if ((wb >> A2B[q]) & 1u) {
            out[q] = a.fb[pop_index<M_F64>(INVA[q], c, g.ncell)];
        } else {
            const double p = a.post[pop_index<M_F64>(q, src, g.ncell)];
            out[q] = in_sponge ? (1.0 - s) * p + s * W64[q] : p;   // feq(q, 1, 0, 0, 0) = w_q
        }
    }
    double sum = 0.0;
#pragma unroll
    for (int q = 0; q < Q; ++q) sum += out[q];
    if (!isfinite(sum)) atomicOr(a.flag, 1);
    if (a.field) a.field[n] = sum;
    if ((m & MK_PROBE) && a.probe_row) {
        int lo = 0, hi = a.n_probe_cells - 1;
        while (lo <= hi) {
            const int mid = (lo + hi) >> 1;
            const long long v = a.probe_cells[mid];
            if (v == c) { a.probe_row[mid] = sum - 1.0; break; }
            if (v < c) lo = mid + 1; else hi = mid - 1;
        }
    }
    if (a.f_out) {
#pragma unroll
        for (int q = 0; q < Q; ++q) __builtin_nontemporal_store(out[q], a.f_out + pop_index<M_F64>(q, c, g.ncell));
        return;
    }
    // the next step: its line-source add (04:229-236), then its collision (01:315-343)
    if (m & MK_LINE) {
        const double v = a.amp * a.line_factor[n] * a.sin_t;
#pragma unroll
        for (int q = 0; q < Q; ++q) out[q] += W64[q] * v * a.dt;
    }
    if (wall) {
#pragma unroll
        for (int q = 0; q < Q; ++q) a.fb_out[pop_index<M_F64>(q, c, g.ncell)] = out[q];
    }
    double sr = 0.0, sux = 0.0, suy = 0.0, suz = 0.0;
#pragma unroll
    for (int p = 0; p < Q; ++p) {
        sr += out[p];
        sux += out[p] * (double)EXA[p];
        suy += out[p] * (double)EYA[p];
        suz += out[p] * (double)EZA[p];
    }
    if (sr == 0.0) sr = 1.0;
    sux /= sr;
    suy /= sr;
    suz /= sr;
    // z periodic (01:283-300): the halo plane 0 collides a copy of plane nz, nz + 1 of plane 1
    const long long zc = k == g.nz ? cell_index(g, i, j, 0) : (k == 1 ? cell_index(g, i, j, g.nz + 1) : -1);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const double fe = feq_a(q, sr, sux, suy, suz);
        const double pv = out[q] - (out[q] - fe) / a.tau;
        __builtin_nontemporal_store(pv, a.post_out + pop_index<M_F64>(q, c, g.ncell));
        if (zc >= 0) a.post_out[pop_index<M_F64>(q, zc, g.ncell)] = pv;
    }
}

hipError_t launch_a_fused(const AFusedArgs &a, hipStream_t s) {
    const long long total = (long long)a.g.nx * a.g.ny * a.g.nz;
    k_a_fused<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(a);
    return hipGetLastError();
}

// Post-collision values of every fluid cell, halo included, of the state f (k_a_collide).
hipError_t launch_a_collide(const double *f, double *post, const uint8_t *mask, const GridDesc &g, double tau,
                            hipStream_t s) {
    const long long all = (long long)(g.nx + 2) * (g.ny + 2) * (g.nz + 2);
    k_a_collide<<<(unsigned)((all + 255) / 256), 256, 0, s>>>(f, post, mask, g, tau);
    return hipGetLastError();
}

// MK_LINE on the listed cells (after clearing it everywhere).
__global__ __launch_bounds__(256) void k_a_mark_line(uint8_t *mask, long long ncell, const long long *cells,
                                                     long long n) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ncell) mask[t] &= (uint8_t)~MK_LINE;
}
__global__ __launch_bounds__(256) void k_a_set_line(uint8_t *mask, const long long *cells, long long n) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) mask[cells[t]] |= MK_LINE;
}

hipError_t launch_a_mark_line(uint8_t *mask, const GridDesc &g, const long long *cells, long long n, hipStream_t s) {
    k_a_mark_line<<<(unsigned)((g.ncell + 255) / 256), 256, 0, s>>>(mask, g.ncell, cells, n);
    if (n > 0) k_a_set_line<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(mask, cells, n);
    return hipGetLastError();
}

// 04_hpc_scalability's apply_source (04:212-235) on the listed fluid cells: the value
// (Amp * spatial) * sin(omega t) (04:229) and the add f += (w_q * value) * dt (04:233-236), in
// the reference's operation order; sin(omega t) and the spatial factors come from the host
// (glibc), so they round as the reference's.
__global__ __launch_bounds__(256) void k_a_line(double *f, GridDesc g, const long long *cells, const double *spatial,
                                                long long n, double amp, double sin_t, double dt) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const long long c = cells[t];
    const double v = amp * spatial[t] * sin_t;
#pragma unroll
    for (int q = 0; q < Q; ++q) f[pop_index<M_F64>(q, c, g.ncell)] += W64[q] * v * dt;
}

hipError_t launch_a_line(double *f, const GridDesc &g, const long long *cells, const double *spatial, long long n,
                         double amp, double sin_t, double dt, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_a_line<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(f, g, cells, spatial, n, amp, sin_t, dt);
    return hipGetLastError();
}

hipError_t launch_a_halo(double *f, const GridDesc &g, hipStream_t s) {
    const long long ring = (2LL * (g.nx + 2) + 2LL * g.ny) * g.nz;
    k_a_halo_xy<<<(unsigned)((ring + 255) / 256), 256, 0, s>>>(f, g);
    const long long plane = (long long)(g.nx + 2) * (g.ny + 2);
    k_a_halo_z<<<(unsigned)((plane + 255) / 256), 256, 0, s>>>(f, g);
    return hipGetLastError();
}

hipError_t launch_a_sweep(const AStepArgs &a, hipStream_t s) {
    const GridDesc &g = a.g;
    const long long all = (long long)(g.nx + 2) * (g.ny + 2) * (g.nz + 2);
    k_a_collide<<<(unsigned)((all + 255) / 256), 256, 0, s>>>(a.f, a.post, a.mask, g, a.tau);
    const long long total = (long long)g.nx * g.ny * g.nz;
    k_a_stream<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace nlos
