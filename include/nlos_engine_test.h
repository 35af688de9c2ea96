/*
 * nlos_engine_test.h - test hooks exported by the engine library beside the drop-in ABI of
 * nlos_engine.h.  They have no counterpart in the reference; parity tests use them to read
 * the full population state and to force a non-finite value from the host.
 */
#ifndef NLOS_ENGINE_TEST_H
#define NLOS_ENGINE_TEST_H

#include "nlos_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Post-collision populations f* of every interior cell as FP64 full populations,
 * array-of-structures [cell][q] over the interior (x fastest); NaN on solid cells. */
int nlos_debug_read_f(nlos_engine *e, double *out);

/* Overwrite the stored population q of interior cell (i,j,k) (0-based) with the full
 * population value f, converted to the engine's storage form. */
int nlos_debug_poke(nlos_engine *e, int32_t i, int32_t j, int32_t k, int32_t q, double f);

/* Plain device copy of `bytes` (16-byte vector loads and stores, one read and one write
 * per byte) repeated `reps` times on a private stream; *avg_ms is the mean launch time
 * from HIP events.  Used to calibrate HBM byte counters and as the practical copy rate. */
int nlos_debug_copy_kernel(size_t bytes, int32_t reps, double *avg_ms);

/* The data movement of the FP16S two-array sweep alone on an nx*ny*nz grid: every interior
 * cell loads the 19 blocked FP16 slots it pulls and stores them into its own slots of a
 * second array (non-temporal), with no mask, bounce-back or collision.  Repeated `reps`
 * times on a private stream; *avg_ms is the mean launch time from HIP events.  The
 * practical ceiling of that access pattern, to set the sweep's time against. */
int nlos_debug_pull_pattern(int32_t nx, int32_t ny, int32_t nz, int32_t reps, double *avg_ms);

#ifdef __cplusplus
}
#endif
#endif /* NLOS_ENGINE_TEST_H */
