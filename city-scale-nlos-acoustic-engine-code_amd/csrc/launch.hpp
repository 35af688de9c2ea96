// Host-side launch interface of the per-storage-mode kernel translation units
// (kernels_f64.hip, kernels_f32.hip, kernels_f16.hip, kernels_b16.hip).  The engine sees only these
// declarations; each mode's TU instantiates Launch<MODE>.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace nlos {

template <int MODE>
struct Launch {
    // Fused collide-stream sweep.  odd: odd-type (LAY_P -> LAY_Q) else even-type.
    static hipError_t step(const StepArgs &a, bool odd, bool capture, dim3 block, hipStream_t s);
    // Fused pull sweep of the two-array pattern (LAY_AB): a.pop -> a.pop_out.
    static hipError_t step_ab(const StepArgs &a, bool capture, dim3 block, hipStream_t s);
    // The same sweep, also recording u of each fluid cell's pulled populations in a.vel
    // (the adjoint's replay; F64 and F32).
    static hipError_t step_ab_vel(const StepArgs &a, dim3 block, hipStream_t s);
    // Ghost layer of the open faces (NSCBC or rest state) written into the current layout.
    static hipError_t ghost(const GhostArgs &a, int lay, hipStream_t s);
    // Pull step of solid cells on open faces (their state feeds the NSCBC ghosts).
    static hipError_t solid(const SolidArgs &a, hipStream_t s);
    // Source injection f += w_q * add on the stencil cells of one step.
    static hipError_t source(const SourceArgs &a, int lay, hipStream_t s);
    // Every population of every cell (halo and padding included) to the rest state.
    static hipError_t fill_rest(void *pop, const GridDesc &g, hipStream_t s);
    // f = feq(rho, 0) on fluid interior cells; rho is interior, x fastest.
    static hipError_t set_density(void *pop, const uint8_t *mask, const GridDesc &g,
                                  const double *rho, float scale, hipStream_t s);
    // Rest state of n tracked solid cells.
    static hipError_t fill_state_rest(void *state, int n, hipStream_t s);
    // Wall copies of LAY_P (see k_step) after the state was set from outside.
    static hipError_t prime_walls(void *pop, const uint8_t *mask, const GridDesc &g, int lay, double alpha,
                                  hipStream_t s);
    // Fill the x-face buffers (StepArgs::fb) from the lattice in layout lay.
    static hipError_t fb_gather(const void *pop, const uint8_t *mask, const GridDesc &g, int lay,
                                void *fb0, void *fb1, float scale, float inv_scale, hipStream_t s);
    // Several scenes of the two-array layout in one launch each.  Argument arrays and the
    // work tables live on the device: the sweep runs a max_tiles x n_scenes grid (n_tiles
    // per scene); blocks[g] = (scene, first item) for the 256-thread workgroups of the
    // ghost and boundary-solid launches.
    static hipError_t step_ab_multi(const StepArgs *scenes, const int *n_tiles, int n_scenes, int max_tiles,
                                    hipStream_t s);
    static hipError_t step_ab_walk(const StepArgs *scenes, const int2 *items, int n_items, int grid, hipStream_t s);
    static hipError_t ghost_multi(const GhostArgs *scenes, const int2 *blocks, int n_blocks, hipStream_t s);
    static hipError_t solid_multi(const SolidArgs *scenes, const int2 *blocks, int n_blocks, hipStream_t s);
    // Source entries of the step named by the device counter; at most max_entries per step.
    static hipError_t source_multi(const SourceMulti &m, long long max_entries, hipStream_t s);
    // Discrete adjoint of one two-array step (see AdjArgs).
    static hipError_t adjoint(const AdjArgs &a, hipStream_t s);
    // Transposed NSCBC ghost fill of every NSCBC face (see AdjArgs), after adjoint().
    static hipError_t adjoint_ghost(const AdjArgs &a, hipStream_t s);
    // out[e] = sum_q w_q lam[q][cells[e]], e < n (FP64).
    static hipError_t source_dot(const void *lam, long long ncell, const long long *cells, int n, double *out,
                                 hipStream_t s);
    // n values of the compute type set to zero.
    static hipError_t fill_compute(void *p, long long n, hipStream_t s);
    // Gather f* of every interior cell into AoS FP64 full populations [cell][q] over the
    // interior, x fastest (for checkpoints and tests).
    static hipError_t gather_f(const void *pop, const uint8_t *mask, const GridDesc &g, int lay,
                               float inv_scale, double *out, hipStream_t s);
};

// Mode-independent helpers (voxel.hip).
hipError_t launch_mask_flags(uint8_t *mask, const GridDesc &g, hipStream_t s);
hipError_t launch_mask_periodic_z(uint8_t *mask, const GridDesc &g, hipStream_t s);
hipError_t launch_fill_f32(float *p, long long n, float v, hipStream_t s);
hipError_t launch_fill_u8(uint8_t *p, long long n, uint8_t v, hipStream_t s);
// Per interior cell: bit q set when c - e_q is solid (two-array bounce-back).
hipError_t launch_wall_bits(const uint8_t *mask, const GridDesc &g, uint32_t *bits, hipStream_t s);
// *ctr += by on the device (one thread).
hipError_t launch_count(long long *ctr, long long by, hipStream_t s);
// Grid-stride copy of n16 16-byte words.
hipError_t launch_copy16(const void *a, void *b, long long n16, hipStream_t s);
// The FP16 two-array sweep's loads and stores alone (kernels_f16.hip); in and out are blocked
// FP16 population arrays of g.ncell cells.
hipError_t launch_pull_pattern16(const void *in, void *out, const GridDesc &g, hipStream_t s);

// Legacy family-A step (legacy_a.hip): the halo of 01's exchange_halos, then collide and
// stream (01's step_pull).
hipError_t launch_a_halo(double *f, const GridDesc &g, hipStream_t s);
hipError_t launch_a_sweep(const AStepArgs &a, hipStream_t s);
// The fused family-A step (see AFusedArgs), the post-collision values of every fluid cell of
// a state, and the MK_LINE bits of the line source's cells.
hipError_t launch_a_fused(const AFusedArgs &a, hipStream_t s);
hipError_t launch_a_collide(const double *f, double *post, const uint8_t *mask, const GridDesc &g, double tau,
                            hipStream_t s);
hipError_t launch_a_mark_line(uint8_t *mask, const GridDesc &g, const long long *cells, long long n, hipStream_t s);
// 04's line source: f_q(cells[e]) += (w_q * ((amp * spatial[e]) * sin_t)) * dt, e < n.
hipError_t launch_a_line(double *f, const GridDesc &g, const long long *cells, const double *spatial, long long n,
                         double amp, double sin_t, double dt, hipStream_t s);

// Voxeliser (voxel.hip).
struct VoxelArgs {
    const double *tri;         // per triangle: v1, v2, v3 (9 doubles), min xy, max xy (4)
    const int *bin_start;      // CSR over gridRows*gridCols bins
    const int *bin_tris;
    int n_tri;
    int grid_rows, grid_cols;
    double geo_min[3], geo_max[3];
    double cell_w, cell_h;
    double x_min, y_min, z_min, dx;
    double sky_z;
    int use_geometry;
    int ground;
};
hipError_t launch_column_heights(const VoxelArgs &v, const GridDesc &g, double *gh, hipStream_t s);
hipError_t launch_mask_from_heights(const VoxelArgs &v, const GridDesc &g, const double *gh,
                                    uint8_t *mask, hipStream_t s);

// Crossing voxeliser (nlos_voxelizer PARITY / WINDING): rays along axis a through the cell
// centres of every interior grid line; the triangles come binned over the two following
// axes (b, c) = (a+1, a+2) mod 3, a rows x cols grid over the mesh's (b, c) box.
struct CrossArgs {
    const double *tri;         // Triangle13: v1, v2, v3 at [0, 9)
    const int *bin_start;      // CSR over rows*cols bins
    const int *bin_tris;
    int rows, cols;
    double bmin, bmax, cmin, cmax, cell_b, cell_c;
    int axis;                  // ray axis a: 0 x, 1 y, 2 z
    int winding;               // 1: signed crossings, 0: crossing parity
    double lo[3], dx;          // domain minimum corner and spacing
    int n[3];                  // interior cells per axis
    double off_b, off_c;       // ray offsets from the cell-centre lattice
    uint8_t *votes;            // interior cells, x fastest: += 1 where solid
    int *overflow;             // set when a line has more crossings than the kernel holds
};
hipError_t launch_crossing_votes(const CrossArgs &a, hipStream_t s);
// Mask from votes: solid where votes >= need (same halo/sky/ground rules as from heights).
hipError_t launch_mask_from_votes(const VoxelArgs &v, const GridDesc &g, const uint8_t *votes, int need,
                                  uint8_t *mask, hipStream_t s);

}  // namespace nlos
