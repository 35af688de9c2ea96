/*
 * nlos_engine.h - C ABI of the MI355X D3Q19 lattice-Boltzmann acoustic engine.
 *
 * Drop-in for the time-stepping path of the reference solver `LBMSolver`
 * (05_sensor_illusion.cpp:241-1122; moving-source variant 08_city_scale_uav.cpp:540-628,
 * 1026-1060).  The reference has no library API: its surface is the LBMSolver class, whose
 * public fields (05:243-268) become `nlos_config` / `nlos_source` / probes, and whose public
 * methods map one to one onto the entry points below:
 *
 *   LBMSolver::init               05:299-450   nlos_create + nlos_load_stl / nlos_set_mask
 *                                              + nlos_set_probes
 *   LBMSolver::apply_source       05:576-667   nlos_apply_source
 *   LBMSolver::exchange_halos     05:671-904   nlos_exchange_halos
 *   LBMSolver::step_collision_bgk 05:906-972   nlos_step_collision
 *   LBMSolver::write_vtk          05:975-1062  nlos_read_field / nlos_write_vtk
 *   LBMSolver::run                05:1064-1121 nlos_step (loop body 05:1077-1093)
 *                                              + nlos_write_probe_files (05:1106-1118)
 *
 * Conventions: every call returns 0 (NLOS_OK) or an nlos_status; nlos_last_error() gives a
 * message for the calling thread.  One host thread per engine.  Work is ordered on the
 * engine's own HIP stream; calls that hand data back to the host synchronise.  Device
 * memory is owned by the engine; every host array is owned by the caller and copied.
 * Non-finite populations are flagged on the device and reported as NLOS_ERR_NONFINITE by
 * the next synchronising call.
 */
#ifndef NLOS_ENGINE_H
#define NLOS_ENGINE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NLOS_ABI_VERSION 1

typedef struct nlos_engine nlos_engine;

typedef enum nlos_status {
    NLOS_OK = 0,
    NLOS_ERR_ARG = 1,         /* bad argument or configuration */
    NLOS_ERR_HIP = 2,         /* HIP runtime failure */
    NLOS_ERR_STATE = 3,       /* call out of order */
    NLOS_ERR_NONFINITE = 4,   /* NaN/Inf density seen during stepping */
    NLOS_ERR_IO = 5,          /* file could not be read or written */
    NLOS_ERR_NOMEM = 6,       /* device allocation failed */
    NLOS_ERR_UNSUPPORTED = 7  /* option not available for this precision */
} nlos_status;

/* Population storage.  NLOS_F64 keeps full populations f in FP64 and follows the
 * reference's arithmetic operation for operation.  NLOS_F32 and NLOS_F16S store the
 * deviation g = f - w_q (rest density 1) in FP32 or in FP16 scaled by 2^f16_scale_log2, and
 * do all collision arithmetic in FP32.  NLOS_B16 stores the same deviations in 16 bits per
 * population as block floating point: a 16-bit mantissa of g_q / w_q per population and one
 * exponent byte per cell (39 bytes per cell instead of FP16's 38), which keeps 15
 * significant bits relative to each cell's largest value; FP32 arithmetic; two-array
 * stream pattern only.  No counterpart in the reference (it stores FP64 only). */
typedef enum nlos_precision { NLOS_F64 = 0, NLOS_F32 = 1, NLOS_F16S = 2, NLOS_B16 = 3 } nlos_precision;

/* Domain faces, used for sponge and NSCBC face sets. */
enum {
    NLOS_FACE_XM = 1, NLOS_FACE_XP = 2, NLOS_FACE_YM = 4,
    NLOS_FACE_YP = 8, NLOS_FACE_ZM = 16, NLOS_FACE_ZP = 32
};
/* The reference's set: every face but the ground (05:491-497 sponge, 05:684-835 NSCBC). */
#define NLOS_FACES_REFERENCE (NLOS_FACE_XM | NLOS_FACE_XP | NLOS_FACE_YM | NLOS_FACE_YP | NLOS_FACE_ZP)

typedef struct nlos_config {
    /* Physical box and spacing; grid = (int)((max-min)/dx + 0.5) per axis (05:332-334). */
    double x_min, x_max, y_min, y_max, z_min, z_max;   /* 05:250-252 */
    double dx;                 /* 05:253 (the programs set 1.5/1.0/0.5 at 05:1129) */
    double c_phys;             /* UnitConverter::c_phys, 340 (05:198); dt = cs*dx/c (05:200-203) */
    double tau;                /* BGK relaxation time, 0.55 (05:268); must exceed 0.5 */
    double sponge_width;       /* metres; 05:455 uses 20*dx */
    double sponge_strength;    /* 0.35 (05:456); sigma = strength*((w-d)/w)^2 (05:498-502) */
    double sky_z;              /* cells above are forced fluid, 120 (05:458, 489) */
    double wall_reflect;       /* 1 = full half-way bounce-back (05:923-926); <1 damps it */
    double world_offset[3];    /* added to the VTK origin when output_world_coords (05:293-296) */
    int32_t sponge_faces;      /* NLOS_FACE_* set, reference NLOS_FACES_REFERENCE */
    int32_t nscbc_faces;       /* NLOS_FACE_* set, reference NLOS_FACES_REFERENCE */
    int32_t ground;            /* 1: bottom halo is solid (05:520-525); 0: open */
    int32_t use_geometry;      /* 05:248 */
    int32_t output_world_coords; /* 05:255 */
    int32_t precision;         /* nlos_precision */
    int32_t f16_scale_log2;    /* NLOS_F16S: stored value = g * 2^k; 0 selects the default 24 */
    int32_t stream_pattern;    /* nlos_pattern */
    int32_t voxelizer;         /* nlos_voxelizer; 0 is the reference's height field */
    int32_t family;            /* nlos_family; 0 is the 05/08 solver */
    int32_t reserved[2];
} nlos_config;

/* Which of the reference's two solvers the engine runs.
 * NLOS_FAMILY_B: 05_sensor_illusion / 08_city_scale_uav (everything above).
 * NLOS_FAMILY_A: the legacy step_pull of 01_acoustic_scattering.cpp and
 *   04_hpc_scalability.cpp, bit for bit in FP64:
 *   - velocity ordering A (01:17-19);
 *   - the source cell collides (BGK, f - (f - feq)/tau);
 *   - a solid source returns the destination's own pre-collision f[inv q];
 *   - the destination's sponge blend follows;
 *   - rho is the sum of the new populations (01:303-359);
 *   - x and y halos at rest, z periodic (01:110-112, 230-301);
 *   - probes at cell (int)((x - min)/dx - 0.5) record rho - 1 (01:448-479);
 *   - dt = sqrt(1/3) dx / c (01:37-41), nz at least 3 (01:100).
 *   The sponge is per cell: by default 01's rule on the x/y faces, with thickness
 *   sponge_width and strength sponge_strength (01:207-225); nlos_set_sponge replaces it
 *   (04's radial ramp).
 *   Needs NLOS_F64 and the two-array pattern.  Its source is 04's line source
 *   (nlos_set_line_source); nlos_set_source, NSCBC, the ground halo, batches and the
 *   adjoint are family-B features and are refused or ignored (ground and NSCBC are off).  nlos_write_vtk writes 01's format (01:414-436, scale 100);
 *   nlos_write_probe_dat writes 01's files including the t = 0 line. */
typedef enum nlos_family { NLOS_FAMILY_B = 0, NLOS_FAMILY_A = 1 } nlos_family;

/* How an STL mesh becomes solid cells.
 * NLOS_VOXEL_HEIGHT is the reference: the top hit of a vertical ray per column, every cell
 * centre below it solid (05:144-192, 469-489).  It fills whatever lies under an overhang.
 * The others decide per cell centre from the mesh crossings of rays through it (no
 * counterpart in the reference):
 *   NLOS_VOXEL_PARITY  - solid where the crossing count from below is odd (closed meshes;
 *                        orientation is not used, overlapping shells cancel);
 *   NLOS_VOXEL_WINDING - solid where the signed crossing count (entering through a face
 *                        whose normal, from the vertex order, points against the ray: +1,
 *                        leaving: -1) is non-zero (outward-oriented meshes; overlapping
 *                        shells give their union, inward shells carve cavities).
 * Adding NLOS_VOXEL_3AXIS casts rays along x, y and z and keeps a cell solid when at least
 * two of them say so, which tolerates holes in the mesh.  Rays sit 1e-7 dx (times an
 * irrational factor) off the cell-centre lattice so that they miss mesh edges placed on it.
 * The mask rules around the mesh (sky_z, halo, ground) are the reference's in every mode;
 * nlos_get_ground_heights returns the column top hits in every mode. */
typedef enum nlos_voxelizer {
    NLOS_VOXEL_HEIGHT = 0, NLOS_VOXEL_PARITY = 1, NLOS_VOXEL_WINDING = 2, NLOS_VOXEL_3AXIS = 4
} nlos_voxelizer;

/* How populations stream.  Both give identical results; they differ in memory and speed.
 * NLOS_PATTERN_INPLACE keeps one array (A-A: alternating own-cell and neighbour sweeps);
 * NLOS_PATTERN_TWO_ARRAY keeps two (pull from one, write the other), which avoids the
 * neighbour writes of the A-A odd sweep.  AUTO takes two arrays when both fit in free
 * device memory. */
typedef enum nlos_pattern {
    NLOS_PATTERN_AUTO = 0, NLOS_PATTERN_INPLACE = 1, NLOS_PATTERN_TWO_ARRAY = 2
} nlos_pattern;

typedef enum nlos_waveform {
    NLOS_WAVE_GAUSS_PULSE = 0, /* amp*exp(-0.5*((t-t0)/sigma_t)^2)   05:579-581 */
    NLOS_WAVE_SINE = 1,        /* amp*sin(2*pi*freq*t)               08:543 */
    NLOS_WAVE_SAMPLED = 2,     /* amp*samples[s] at step s (s = round(t/dt)) */
    NLOS_WAVE_RICKER = 3       /* amp*(1-2a)exp(-a), a=(pi*freq*(t-t0))^2 */
} nlos_waveform;

typedef enum nlos_injection {
    NLOS_INJECT_GAUSS5 = 0,    /* normalised Gaussian stencil, R=2, sigma=1.25 cells (05:587-665) */
    NLOS_INJECT_TRILINEAR = 1  /* 8 cell-centred trilinear weights, normalised over fluid cells */
} nlos_injection;

typedef struct nlos_source {
    double sx, sy, sz;         /* DroneSource start (05:213) */
    double ex, ey, ez;         /* DroneSource end (05:214) */
    double speed;              /* m/s; <1e-12 keeps the source at the start (05:225) */
    double amp;                /* 05:216 */
    double freq;               /* Hz (08:543) */
    double t0, sigma_t;        /* pulse centre/width, 0.05 / 0.005 (05:579-580) */
    double gain;               /* source_gain, 1.0 (05:583) */
    double stencil_sigma_cells;/* 1.25 (05:588) */
    int32_t stencil_radius;    /* 2 (05:587) */
    int32_t waveform;          /* nlos_waveform */
    int32_t injection;         /* nlos_injection */
    int32_t n_samples;         /* NLOS_WAVE_SAMPLED: number of samples */
    const double *samples;     /* NLOS_WAVE_SAMPLED: copied by nlos_set_source */
} nlos_source;

typedef struct nlos_info {
    int32_t nx, ny, nz;        /* interior cells NXg, NYg, NZg (05:332-334) */
    int32_t precision;
    double dt;                 /* seconds (05:200-203) */
    int32_t stream_pattern;    /* the pattern in use (NLOS_PATTERN_INPLACE or _TWO_ARRAY) */
    int32_t reserved_info;
    int64_t step;              /* next time index s; t = s*dt (05:1079) */
    int64_t n_probes;
    int64_t n_solid_boundary;  /* solid cells on NSCBC faces (their state is tracked) */
    double bytes_per_update;   /* algorithmic HBM bytes per lattice update of the step kernel */
    size_t device_bytes;       /* device memory held by the engine */
} nlos_info;

typedef struct nlos_timing {
    int64_t launches;          /* step-kernel launches timed since nlos_timing_reset */
    double total_ms;           /* summed step-kernel durations (HIP events on the engine stream) */
    double avg_ms;
} nlos_timing;

/* ---- lifetime ---------------------------------------------------------------------- */
int nlos_abi_version(void);
const char *nlos_last_error(void);
/* Fills the reference defaults of 05:243-268, 455-458, 669 (DX=2, box 1650x1150x250 m). */
int nlos_config_default(nlos_config *cfg);
/* Fills the reference defaults of 05:579-589 (pulse t0=0.05, sigma_t=0.005, R=2, 1.25). */
int nlos_source_default(nlos_source *src);
/* LBMSolver::init (05:299-395): grid, tables, f = w everywhere, empty geometry. */
int nlos_create(const nlos_config *cfg, nlos_engine **out);
int nlos_destroy(nlos_engine *e);
int nlos_get_info(const nlos_engine *e, nlos_info *info);

/* ---- geometry ------------------------------------------------------------------------ */
/* GeometryEngine::load + buildGrid (05:70-142) on the host, then the column top-hit ray
 * kernel (getGroundHeight, 05:144-192) and the mask rule of init_domain (05:453-532) on
 * the GPU.  A file that cannot be read returns NLOS_ERR_IO and leaves empty geometry, as
 * the reference warns and continues (05:407-410). */
int nlos_load_stl(nlos_engine *e, const char *path, double scale, double sx, double sy, double sz);
int nlos_voxelize_stl(nlos_engine *e, const void *stl, size_t len,
                      double scale, double sx, double sy, double sz);
/* Column heights gH of the last voxelisation over (nx+2)*(ny+2) columns incl. halo. */
int nlos_get_ground_heights(nlos_engine *e, double *out);
/* Interior solid mask, nx*ny*nz bytes, x fastest, 0 = FLUID, 1 = SOLID (05:36). */
int nlos_set_mask(nlos_engine *e, const uint8_t *mask);
int nlos_get_mask(nlos_engine *e, uint8_t *mask);

/* ---- initial state, source, probes --------------------------------------------------- */
/* NLOS_FAMILY_A only: sponge strength per interior cell, x fastest (replaces 01's default
 * rule; 04:176-208 is a radial one).  A cell blends when sigma > 1e-6 (01:313). */
int nlos_set_sponge(nlos_engine *e, const double *sigma);
/* NLOS_FAMILY_A only: the line source of 04_hpc_scalability.cpp (apply_source, 04:212-235),
 * added to the populations at the start of every step, before the halo exchange
 * (04:369-388).  Every fluid interior cell, in every z layer, whose centre
 * (x, y) = (x_min + i dx + dx/2, y_min + j dx + dx/2) has d2 = (x - x0)^2 + (y - y0)^2 <= r2_max
 * receives f_q += w_q v dt with v = amp exp(-ln 2 d2 / b) sin(omega t), t = (s - step0) dt at
 * step s.  04's values: x0 = 0, y0 = -6, amp = 0.01, b = 2, omega = pi/15, r2_max = 36,
 * dt = 0.577350269 dx / c (04:32-37, 65-72); its timed loop restarts t at 0 after 100 warm-up
 * steps (step0 = 100 then).  Setting it again replaces it; amp = 0 switches it off. */
typedef struct nlos_line_source {
    double x0, y0;
    double amp, b, omega, r2_max, dt;
    int64_t step0;
} nlos_line_source;
int nlos_set_line_source(nlos_engine *e, const nlos_line_source *s);
/* Initial populations f = feq(rho, 0) = w_q*rho on fluid cells (interior rho, x fastest). */
int nlos_set_density(nlos_engine *e, const double *rho);
int nlos_set_source(nlos_engine *e, const nlos_source *src);
/* Probe positions (x,y,z triples, metres); cell = floor((p-min)/dx) (05:417-419).  A probe
 * outside the grid is NLOS_ERR_ARG; one inside a solid is accepted and records 0 (rho=1,
 * 05:936) - its flag is set in in_solid (may be NULL). */
int nlos_set_probes(nlos_engine *e, const double *xyz, int32_t n, uint8_t *in_solid);

/* ---- the time step -------------------------------------------------------------------- */
/* The three phases of one reference step, for callers that drive them one by one. */
int nlos_apply_source(nlos_engine *e, double t);   /* 05:576-667 */
int nlos_exchange_halos(nlos_engine *e);           /* 05:671-904 (single domain: ghosts only) */
int nlos_step_collision(nlos_engine *e);           /* 05:906-972 + probe record 05:1085-1093 */
/* n fused steps s = step .. step+n-1, each apply_source(s*dt), exchange_halos,
 * step_collision, probe record.  flags: NLOS_STEP_CAPTURE_FIELD keeps rho-1 of the last
 * step for nlos_read_field.  Two-array engines replay a captured HIP graph of two steps
 * (captured on the first call and again when a buffer it uses grows); family-A engines run
 * a step as one fused pass.  Both are bit-identical to the separate launches. */
#define NLOS_STEP_CAPTURE_FIELD 1
int nlos_step(nlos_engine *e, int64_t n, int32_t flags);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * nlos_step_noalloc runs the same n steps as nlos_step, bit for bit, and calls no
 * hipMalloc, hipFree or other allocating routine.  The per-step source schedule goes through
 * one static device array per GPU of NLOS_NOALLOC_SOURCE_ENTRIES entries (4 MiB, held until
 * the process ends) instead of the engine's growable buffers.  It refuses (engine
 * unchanged) what would need an allocation: a probe history past its capacity
 * (nlos_set_probes reserves NLOS_PROBE_HISTORY_RESERVE steps, nlos_step grows it),
 * NLOS_STEP_CAPTURE_FIELD before the first capture by nlos_step, timing enabled, a family-A
 * engine, or an engine whose device is not current.  Calls on one GPU share its array: two
 * engines stepped this way on different streams at the same time race. */
#define NLOS_NOALLOC_SOURCE_ENTRIES 262144
#define NLOS_PROBE_HISTORY_RESERVE 4096
int nlos_step_noalloc(nlos_engine *e, int64_t n, int32_t flags);
int nlos_synchronize(nlos_engine *e);

/* ---- results -------------------------------------------------------------------------- */
/* Probe histories rho-1, probe-major [probe][step], steps recorded so far (at most
 * max_steps per probe).  *n_steps receives the number recorded. */
int nlos_read_probes(nlos_engine *e, double *out, int64_t max_steps, int64_t *n_steps);
/* float(rho-1) of the last captured step over the interior, x fastest (05:979-985). */
int nlos_read_field(nlos_engine *e, float *out);
/* probe_<name>.txt, "%.8e" one value per line (05:1110-1114). */
int nlos_write_probe_files(nlos_engine *e, const char *dir, const char *const *names);
/* Legacy ASCII VTK of the captured field, byte-compatible with 05:1045-1053. */
int nlos_write_vtk(nlos_engine *e, const char *path);

/* VTK encodings of the captured field.  NLOS_VTK_ASCII is the reference's text
 * (05:1045-1053); NLOS_VTK_BINARY is the legacy VTK BINARY form of the same dataset: the
 * header with "BINARY" in place of "ASCII", then the nx*ny*nz float32 values big-endian,
 * x fastest.  No counterpart in the reference (it writes ASCII only). */
#define NLOS_VTK_ASCII 0
#define NLOS_VTK_BINARY 1
int nlos_write_vtk_binary(nlos_engine *e, const char *path);

/* Field dump off the step stream: copies the captured field to pinned host memory on a
 * second stream and writes the file (NLOS_VTK_ASCII or NLOS_VTK_BINARY) on a host thread,
 * returning at once.  Steps keep running meanwhile; a later step that captures the field
 * waits on the device for the copy to finish.  One dump is in flight per engine (a second
 * call first waits for the first).  nlos_dump_wait waits for it and returns its status; the
 * reference writes frames synchronously inside run (05:1095-1103). */
int nlos_dump_field_async(nlos_engine *e, const char *path, int32_t format);
int nlos_dump_wait(nlos_engine *e);

/* Probe histories in the family-A text format of 01_acoustic_scattering.cpp:481-495:
 * dir/probe_<name>.dat, a "# Time Pressure" line, then one "%g %g" line per recorded step,
 * time = s*dt after s steps (s = 1, 2, ...) and pressure = scale*(rho-1) (01:466, whose
 * output_scale_factor is 100).  The probe cells are the engine's (05:415-437). */
int nlos_write_probe_dat(nlos_engine *e, const char *dir, const char *const *names, double scale);

/* ---- checkpoint / restore ------------------------------------------------------------ */
/* Serialises populations, tracked boundary-solid state, step index and probe histories.
 * With buf == NULL only *needed is set.  Halo (ghost) values are not part of the state: a
 * caller that drives the phases one by one calls nlos_exchange_halos after nlos_restore,
 * before nlos_step_collision, as every reference step does (nlos_step always does). */
int nlos_checkpoint(nlos_engine *e, void *buf, size_t cap, size_t *needed);
int nlos_restore(nlos_engine *e, const void *buf, size_t len);

/* ---- measurement ------------------------------------------------------------------------ */
/* When on, every step-kernel launch is bracketed by HIP events on the engine stream. */
int nlos_timing_enable(nlos_engine *e, int32_t on);
int nlos_timing_reset(nlos_engine *e);
int nlos_timing_get(nlos_engine *e, nlos_timing *t);

/* ---- gradients (discrete adjoint) ------------------------------------------------------
 * No counterpart in the reference, which has no differentiation.  The adjoint transposes
 * the two-array step exactly (pull with half-way bounce-back and wall reflection, BGK
 * collision, sponge blend, probe record, source injection) and the NSCBC ghost fill of the
 * open faces (05:684-835); faces without NSCBC hold the rest state.  It needs
 * stream_pattern NLOS_PATTERN_TWO_ARRAY, precision NLOS_F64 or NLOS_F32, and, with NSCBC
 * faces, no solid cell on an NSCBC boundary plane (their tracked state is not transposed;
 * NLOS_ERR_UNSUPPORTED).
 *
 * nlos_record_forward runs n steps like nlos_step and keeps the state at its first step and
 * every `every` steps after (device memory, one population array each).
 * nlos_backward then returns the gradient of a loss L over the recorded steps, given
 * dL/d(rho-1) for every probe and step ([probe][step], stride max_steps, global step
 * indices as in nlos_read_probes).  It recomputes each segment of `every` steps from its
 * checkpoint (every + 1 population arrays and 3 values per cell and step of device
 * memory), runs the adjoint step by step
 * back to the segment start, and accumulates in FP64:
 *   d_position: dL/d(shift of the whole source path, start and end together), per metre;
 *   d_amp:      dL/d amp;
 *   d_samples:  for NLOS_WAVE_SAMPLED, dL/d samples[k] (caller array of n_samples, or NULL).
 * The stencil cell set of the injection is piecewise constant in the position; the
 * gradient is the derivative within the current set.  The engine state is not changed. */
typedef struct nlos_grad {
    double d_position[3];
    double d_amp;
    double *d_samples;
    int64_t n_samples;
} nlos_grad;

int nlos_record_forward(nlos_engine *e, int64_t n, int32_t every);
int nlos_backward(nlos_engine *e, const double *dl_dprobe, int64_t max_steps, nlos_grad *grad);

/* ---- batched scenes ------------------------------------------------------------------
 * No counterpart in the reference, which steps one domain per process (05:1168-1170).  A
 * batch owns n independent scenes, each an nlos_engine of its own configuration (same
 * precision for all; two-array stream pattern) that the caller sets up through the
 * ordinary calls (nlos_set_mask, nlos_set_source, nlos_set_probes, nlos_set_density ...)
 * and reads back through them (nlos_read_probes ...).  nlos_batch_step advances every scene
 * together: per step one source, one ghost, one boundary-solid and one sweep launch
 * cover all scenes, and pairs of steps replay one captured HIP graph.  A scene of a
 * batch cannot be stepped, restored or destroyed on its own (NLOS_ERR_STATE).  Each scene's
 * results are bit-identical to the same scene stepped alone. */
typedef struct nlos_batch nlos_batch;

typedef struct nlos_tune {
    int32_t block_x, block_y;   /* one tile shape for every scene (block_x * block_y = 256),
                                   or 0, 0 when per_scene is set */
    int32_t candidates;         /* sweeps timed */
    int32_t per_scene;          /* 1: each class of row length has its own tile width
                                   (nlos_batch_tile_widths) */
    double best_ms;             /* sweep time of the choice, all scenes, one step */
} nlos_tune;

int nlos_batch_create(const nlos_config *cfgs, int32_t n, nlos_batch **out);
int nlos_batch_destroy(nlos_batch *b);
int nlos_batch_size(const nlos_batch *b, int32_t *n);
/* The engine handle of scene i (owned by the batch). */
int nlos_batch_scene(nlos_batch *b, int32_t i, nlos_engine **scene);
/* Picks the sweep tile shapes: tiles of w x 256/w cells, w = 64, 128 or 256.  Times the
 * sweep of all scenes with each w, then each class of row length (interior rows of 1, 2, 3,
 * or more 64-cell waves) alone with each w, then all scenes with every class at its best w,
 * which is kept when it is at least 1 % faster than the best single w (each timing the
 * fastest of reps launches, into the array the next step overwrites, so the state is
 * untouched).  result may be NULL. */
int nlos_batch_autotune(nlos_batch *b, int32_t reps, nlos_tune *result);
/* The sweep tile width of every scene (n = the batch size), and setting them (each 64, 128
 * or 256), e.g. to reuse an autotuned choice. */
int nlos_batch_tile_widths(const nlos_batch *b, int32_t *widths, int32_t n);
int nlos_batch_set_tile_widths(nlos_batch *b, const int32_t *widths, int32_t n);
/* n steps of every scene (each: apply_source(s*dt), exchange_halos, step_collision, probe
 * record, as nlos_step).  Each step is one launch per kernel over all scenes; the sweep's is
 * a persistent grid of workgroups walking the (scene, tile) work list. */
int nlos_batch_step(nlos_batch *b, int64_t n);
int nlos_batch_synchronize(nlos_batch *b);
/* When on, each nlos_batch_step call is bracketed by HIP events on the batch stream;
 * nlos_timing.launches counts steps and avg_ms is the time per step of all scenes. */
int nlos_batch_timing_enable(nlos_batch *b, int32_t on);
int nlos_batch_timing_get(nlos_batch *b, nlos_timing *t);

#ifdef __cplusplus
}
#endif
#endif /* NLOS_ENGINE_H */
