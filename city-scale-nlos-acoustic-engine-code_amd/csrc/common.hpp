// Shared definitions of the D3Q19 engine: lattice tables, grid description, mask bits,
// kernel argument blocks.  Included by host and device translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nlos {

constexpr int Q = 19;

// Velocity ordering B of the reference (05_sensor_illusion.cpp:21-23).
constexpr int EX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1, 0, 0, 0, 0};
constexpr int EY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1};
constexpr int EZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
// inv_q as the search of 05:301-312 yields it: e_INV[q] = -e_q.
constexpr int INV[Q] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};

// Weights (05:25-31).
constexpr double W64[Q] = {
    1.0 / 3.0,
    1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};

constexpr double CS2 = 1.0 / 3.0;                  // 05:17
constexpr double CS = 0.57735026918962576451;      // 05:18
constexpr double PI = 3.14159265358979323846;      // 05:19

// Storage modes.  F64: full populations f, FP64, reference arithmetic order.  F32/F16:
// deviations g = f - w_q, FP32 arithmetic, stored in FP32 or in FP16 times 2^k.  B16:
// block floating point per cell in normalised units x_q = g_q / w_q.  The rest slot x_0,
// which only its own cell reads, is FP32; slots 1..18 hold 16-bit signed mantissas m_q and
// one signed byte E_c per cell is their shared exponent: x_q = m_q * 2^E_c.  The mantissas
// of a cell are scaled so that the largest |x_q| fills 15 bits, which keeps 15 significant
// bits relative to the cell's own amplitude instead of FP16's 11 per value (41 bytes per
// cell; layout at b16_rest in kernels.hpp).  The exponent ties the 18 slots of a cell
// together, so B16 runs only with the two-array pattern, where every cell has one writer
// per step.  B16 mantissas have no NaN or Inf: a non-finite density is reported by the
// sweep of the step in which it forms (StepArgs::flag).
enum Mode : int { M_F64 = 0, M_F32 = 1, M_F16 = 2, M_B16 = 3 };
constexpr int B16_BITS = 15;        // mantissa magnitude bits
constexpr int B16_EMIN = -113;      // smallest exponent e of max|g/w| < 2^e; E_c = e - 15 >= -128

// Mask byte bits.  The array covers the halo: ground halo cells are SOLID, the other halo
// cells are fluid (05:509-531).
constexpr uint8_t MK_SOLID = 1;   // CellType SOLID (05:36)
constexpr uint8_t MK_NEAR = 2;    // fluid cell with a solid cell among its 18 neighbours
constexpr uint8_t MK_PROBE = 4;   // fluid or solid cell that carries a probe

// Face bits as in nlos_engine.h.
constexpr int F_XM = 1, F_XP = 2, F_YM = 4, F_YP = 8, F_ZM = 16, F_ZP = 32;

// Grid: interior nx*ny*nz, one halo layer, x pitch px >= nx+2 (a multiple of 64).  Cell
// (i,j,k) with i in [0,nx+1] etc. sits at origin + i + px*(j + py*k); population q of it
// at pop_index<MODE>(q, cell, ncell).  origin = 63 puts every interior row start (i = 1) on a 64-element
// boundary, so a wavefront's 64 consecutive cells fill whole cache lines.
struct GridDesc {
    int nx, ny, nz;
    int px, py, pz;
    long long plane;   // px*py
    long long ncell;   // population stride, a multiple of 64
    long long origin;
};

__host__ __device__ __forceinline__ long long cell_index(const GridDesc &g, int i, int j, int k) {
    return g.origin + (long long)i + (long long)g.px * ((long long)j + (long long)g.py * k);
}

// Population arrays of the 16-bit modes (FP16, B16) are blocked by 64 consecutive cells (one
// wavefront's row segment): a block holds the 64 values of slot 0, then the 64 of slot 1, and
// so on (array of structures of arrays).  The Q slots of a block are one contiguous run of
// memory, so the sweep streams one run per block instead of Q runs a whole array apart.
// FP32 and FP64 arrays stay slot-major (slot q of cell c at q*ncell + c).  Measured on the
// two-array sweep at 1024x1024x256: FP16 4.08 -> 3.87 ms and B16 4.74 -> 4.52 ms blocked;
// FP32 7.84 -> 7.82 (no gain) and FP64 13.70 -> 15.40 (a loss), where a wave's run of one
// slot is already 256 or 512 bytes.  ncell is a multiple of 64, so a blocked index of any
// cell in [0, ncell) lies in [0, Q*ncell).
constexpr int PBLK = 64;
__host__ __device__ constexpr bool blocked_mode(int mode) { return mode == M_F16 || mode == M_B16; }
template <int MODE>
__host__ __device__ __forceinline__ long long pop_index(int q, long long c, long long ncell) {
    if constexpr (blocked_mode(MODE)) return ((c >> 6) * Q + q) * PBLK + (c & (PBLK - 1));
    else return (long long)q * ncell + c;
}
inline long long pop_index_host(int mode, int q, long long c, long long ncell) {
    return blocked_mode(mode) ? pop_index<M_F16>(q, c, ncell) : pop_index<M_F64>(q, c, ncell);
}
// B16 mantissa arrays hold slots 1..18 only: the same blocking with Q - 1 slots.
__host__ __device__ __forceinline__ long long mant_index(int q, long long c) {
    return ((c >> 6) * (Q - 1) + (q - 1)) * PBLK + (c & (PBLK - 1));
}

// Lattice layouts of the in-place (A-A) scheme.  Post-collision population q of cell x:
//   LAY_P (after an even-type sweep, and initially): slot INV[q] of x itself;
//   LAY_Q (after an odd-type sweep): slot q of x+e_q, or slot INV[q] of x when x+e_q is
//   solid (the half-way bounce-back is folded into the store).
// An odd-type sweep turns P into Q, an even-type sweep Q into P.  Solid cells are never
// swept; in LAY_P the slot q of a wall cell w holds the reflected f*_q(w - e_q) of its
// fluid neighbour, so that the odd sweep reads bounced populations like streamed ones.
//   LAY_AB (two arrays, pull sweep): slot q of x in the current array; the other array
//   receives the next step.  Bounce-back reads the cell's own slot INV[q]; the slots of
//   solid cells hold the rest state and are never read.
enum Layout : int { LAY_P = 0, LAY_Q = 1, LAY_AB = 2 };

struct StepArgs {
    void *pop;
    void *pop_out;                       // LAY_AB: array receiving the step (else unused)
    const uint8_t *mask;
    const void *sig_x, *sig_y, *sig_z;   // per-axis sponge tables (compute type)
    GridDesc g;
    double omega;
    double alpha;                        // wall reflection, 1 = reference
    float scale, inv_scale;              // F16 storage scale and its inverse
    const long long *probe_cells;        // sorted unique probe cells
    int n_probe_cells;
    double *probe_row;                   // this step's row of the probe history
    const long long *step_ctr;           // if set: the row is hist + (*step_ctr - ctr_lag) * n_probe_cells
    int ctr_lag;                         // 1 when the counter was advanced before the sweep
    double *hist;
    float *field;                        // interior rho-1 (capture sweeps only)
    int *flag;                           // bit 0: non-finite density seen
    int tiles_x, tiles_y;                // block tiling of the interior
    int nblocks;
    int dpp;                             // every lane active: odd sweep may shift across lanes
    int tile_w;                          // multi-scene sweep: tile width 64, 128 or 256 (tiles
                                         // of tile_w x 256 / tile_w cells; tiles_x, tiles_y match)
    void *fb[2];                         // x-face buffers (x-, x+) or null, see below
    const uint32_t *wall_bits;           // two-array sweep: bit q set when c - e_q is solid
    const void *xh;                      // x-halo records or null, see below
    void *vel;                           // velocity sweeps (the adjoint's replay): u of f_in,
                                         // [3][ncell] in the compute type
};

// x-face buffers.  The cells I = (1, j, k) and (nx, j, k) feed the NSCBC ghosts of the x
// faces, but in the x-fastest layout their populations sit one per cache line.  The sweep
// therefore also writes the post-collision populations of those cells, in storage form,
// to fb[face][(j-1) + ny*(k-1)][q]; the source kernel adds to them like to the lattice, and
// the ghost kernel reads them contiguously.  B16 records are FP32 deviations g equal to the
// decoded stored values.

// x-halo records (two-array pattern).  The halo cells i = 0 and i = nx + 1 sit alone in
// their cache lines (their 64-cell blocks in the 16-bit modes), so storing the ghost values
// there and pulling them back moves a whole line per slot for one value each way: 85 of
// the FP16 ghost kernel's 111 us per step at 1024x1024x256 went to those stores.  When xh
// is set, the five slots that interior cells pull from such a cell (EX = +1 at i = 0,
// EX = -1 at i = nx + 1) are kept in xh instead, in storage form, rows (j, k) over the whole
// halo ring [0, ny+1] x [0, nz+1]: slot q of the cell at element xh_index of the storage
// type (B16: the 16-bit mantissa, with the cell's exponent byte at xh_exp_offset).  The
// ghost kernel writes them and the sweep and the solid kernel read them, within one step;
// the lattice slots of those cells are not written and not read.
// Rank of slot q among the five slots with the same nonzero EX[q] (-1 where EX[q] = 0).
constexpr int XR[Q] = {-1, 0, 0, -1, -1, -1, -1, 1, 2, 1, 2, 3, 4, 3, 4, -1, -1, -1, -1};
__host__ __device__ __forceinline__ long long xh_rows(const GridDesc &g) {
    return (long long)(g.ny + 2) * (g.nz + 2);
}
// slot q (EX[q] != 0) of the halo cell (0, j, k) if EX[q] = +1, (nx + 1, j, k) if EX[q] = -1
__host__ __device__ __forceinline__ long long xh_index(const GridDesc &g, int q, int j, int k) {
    return (long long)((EX[q] > 0 ? 0 : 5) + XR[q]) * xh_rows(g) + j + (long long)(g.ny + 2) * k;
}
// byte offset of the B16 exponent of the halo cell on face f (0: i = 0, 1: i = nx + 1)
__host__ __device__ __forceinline__ long long xh_exp_offset(const GridDesc &g, int f, int j, int k) {
    return 20 * xh_rows(g) + f * xh_rows(g) + j + (long long)(g.ny + 2) * k;
}
// bytes of the record buffer for elements of s_bytes bytes
__host__ __device__ __forceinline__ long long xh_bytes(const GridDesc &g, int mode, int s_bytes) {
    return 10 * s_bytes * xh_rows(g) + (mode == M_B16 ? 2 * xh_rows(g) : 0);
}

struct GhostArgs {
    void *pop;
    const uint8_t *mask;
    GridDesc g;
    const void *fb[2];                    // x-face buffers (see StepArgs)
    int nscbc_faces;
    const void *s_old;                    // tracked state of solid cells on open faces
    const int *face_map[6];               // per NSCBC face: plane index -> state slot or -1
    float scale, inv_scale;
    long long n_halo;                     // ghost_items(g, skip_zm)
    void *xh;                             // x-halo records or null (see StepArgs)
    int skip_zm;                          // the k = 0 layer is solid ground: not enumerated
    long long *bump;                      // if set: one thread adds 1 (a device step counter)
};

// Lanes per halo row of the ghost kernel: the row's cells i in [1, nx], 64-cell aligned.
__host__ __device__ __forceinline__ long long ghost_row_lanes(int nx) { return ((long long)nx + 63) / 64 * 64; }
// Work items of the ghost kernel: the rows of the z layers (k = nz + 1, and k = 0 unless
// skip_zm) and the two y rows of every interior k, at ghost_row_lanes each plus their two
// end cells (i = 0, nx + 1), then one item per x-halo cell of the interior rows.
__host__ __device__ __forceinline__ long long ghost_items(const GridDesc &g, int skip_zm) {
    const long long rows = (long long)(2 - (skip_zm ? 1 : 0)) * (g.ny + 2) + 2LL * g.nz;
    return rows * (ghost_row_lanes(g.nx) + 2) + 2LL * g.ny * g.nz;
}

struct SolidArgs {
    void *pop;
    const uint8_t *mask;
    GridDesc g;
    int lay;
    const long long *cells;               // boundary solid cells
    int n;
    const void *s_old;
    void *s_new;
    double alpha;
    float scale, inv_scale;
    const void *xh;                       // x-halo records or null (see StepArgs)
};

struct SourceArgs {
    void *pop;
    const uint8_t *mask;
    GridDesc g;
    void *fb[2];                          // x-face buffers (see StepArgs)
    const long long *cells;               // fluid cells of this step's stencil
    const double *adds;                   // A*ww per cell (05:660)
    int n;
    double alpha;                         // wall reflection (for the wall copies)
    float scale, inv_scale;
};

// Discrete adjoint of one two-array step.  state holds S'_n, the populations the step
// pulled from (after the source add, with the ghost values of that step in the halo);
// lam_in the adjoint of S_{n+1} (what the step wrote), lam_out receives the adjoint of
// S'_n.  Both adjoint arrays are [q][ncell] in the compute type.  On NSCBC faces
// (nscbc_faces) the adjoint of a pulled ghost slot is left in that halo slot of lam_out;
// the ghost pass (k_adjoint_ghost, one thread per boundary cell) then moves it to the
// boundary cell the ghost was computed from and clears it.
struct AdjArgs {
    const void *state;
    const void *lam_in;
    void *lam_out;
    const uint8_t *mask;
    const uint32_t *wall_bits;
    const void *sig_x, *sig_y, *sig_z;
    GridDesc g;
    double omega;
    double alpha;
    float scale, inv_scale;
    const long long *probe_cells;         // sorted unique probe cells
    int n_probe_cells;
    const double *dl_dp;                  // dL/d(rho-1) of this step per unique probe cell, or null
    int nscbc_faces;                      // faces whose halo holds NSCBC ghosts (F_XM ...)
    const void *vel;                      // u of the step's pulled populations, [3][ncell]
};

// One step of the legacy family-A mode (legacy_a.hip).  f holds the populations in
// ordering A and is updated in place; post receives the post-collision values.
struct AStepArgs {
    double *f;
    double *post;
    const uint8_t *mask;
    const double *sigma;                  // sponge per interior cell, x fastest
    GridDesc g;
    double tau;
    const long long *probe_cells;         // sorted unique probe cells
    int n_probe_cells;
    double *probe_row;                    // rho - 1 after this step per unique probe cell
    double *field;                        // rho per interior cell (capture steps) or null
    int *flag;                            // bit 0: non-finite density seen
};

// One family-A step in one pass (legacy_a.hip, k_a_fused).  The stream of the step pulls
// from `post`, the post-collision values of the step's start state (every fluid cell, halo
// cells included), and bounces from `fb`, that start state after its line-source add, read
// only at cells next to a wall.  The last step of a run writes the new state to f_out;
// every other step applies the next step's line-source add to the new state and collides it
// right away: post_out receives the next step's post-collision values (planes nz and 1 also
// into their periodic z-halo planes 0 and nz + 1), fb_out the next bounce sources.
constexpr uint8_t MK_LINE = 8;            // family A: fluid cell of the line source (line_factor)
struct AFusedArgs {
    const uint32_t *wall_bits;            // per cell: bit b set when c - e_b (ordering B) is solid
    const double *post;
    const double *fb;
    double *f_out;                        // last step (else null)
    double *post_out, *fb_out;            // other steps
    const uint8_t *mask;
    const double *sigma;                  // sponge per interior cell, x fastest
    const double *line_factor;            // exp(-ln2 d^2 / b) per interior cell, x fastest (MK_LINE)
    double amp, sin_t, dt;                // line source of the next step: f += (w (amp s) sin_t) dt
    GridDesc g;
    double tau;
    const long long *probe_cells;
    int n_probe_cells;
    double *probe_row;
    double *field;
    int *flag;
};

// Source entries of several scenes over a run of steps, read at the step the device counter
// names: entries step_off[t] .. step_off[t+1] of step t = *ctr - *ctr_base.
struct SourceMulti {
    const SourceArgs *scenes;
    const long long *ctr, *ctr_base;
    const long long *step_off;
    const int *entry_scene;               // scene of each entry (null: all of scene 0)
    const long long *cells;
    const double *adds;
};

}  // namespace nlos
