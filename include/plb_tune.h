/*
 * plb_tune.h -- calibration and autotuning for libplb.  These entry points
 * have no counterpart in the reference (it has no GPU code); they serve the
 * benchmark, the command lines and callers that pick a launch shape.
 */
#ifndef PLB_TUNE_H
#define PLB_TUNE_H

#include "plb.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Streaming copy of `bytes` bytes (16-byte vectors, one read and one write
 * per element) on `device`, best of `iters` launches after one warm-up.
 * *gbs = 2*bytes / kernel time: the practical HBM ceiling the step kernel is
 * compared with.  Allocates 2*bytes of device memory for the call. */
int plb_copy_bandwidth(int device, int64_t bytes, int iters, double *gbs);
/* bytes must be below 2^39 (one 16-byte element per thread of a 1-D grid). */

/* One timed candidate of the autotuner. */
typedef struct {
    plb_config cfg;
    double mlups;   /* pore-node updates per microsecond over `steps` steps */
    double ms;      /* device time of `steps` steps */
} plb_tune_result;

/* Times candidate launch shapes for `base` (precision, bounce-back, tau, gx
 * kept) on the mask, `steps` timed steps each, and writes the fastest
 * configuration to *best.  Stage 1: index encoding (I32, TILE8, and BITMAP
 * for half-way bounce-back) x layout (SOA, AOSOA) x node order (reference;
 * 8^3 bricks; pencils of 64 and 128 full rows) at 256 threads, each on a
 * context of its own.  Stage 2: block size (128, 256, 512) x nodes per thread
 * (1, 2; 1 only for AA streaming) x waves per SIMD (no cap, 6, 5, 4) for the
 * stage-1 winner, all timed on one context, since they change only the
 * launch.  The stream mode of base is kept.  All candidates give
 * bit-identical results.  When cache_path
 * is not NULL, a result stored there for the same mask (64-bit hash of the
 * voxels and dimensions), precision and bounce-back mode is returned without
 * timing, and new results are appended (one text line per result).
 * results (may be NULL) receives up to max_results timed candidates. */
int plb_autotune(const plb_config *base, const uint8_t *flag, int nx, int ny, int nz, int num_p,
                 int steps, const char *cache_path, plb_config *best, plb_tune_result *results,
                 int max_results, int *n_results);

/* Launch-shape search for a batch of samples (plb_batch_create's inputs):
 * times `steps` fixed steps per candidate with plb_batch_run, or with
 * plb_batch_run_persistent when `persistent` is nonzero, and writes the
 * fastest configuration to *best (precision, bounce-back and nodes per
 * thread 1 kept from base).  Stage 1: index encoding (I32, TILE8) x layout (SOA, AOSOA) x
 * node order (V order; pencils of 16, 32 and 64 full rows of each sample) at
 * base->block_size, one batch each.  Stage 2: block size 128, 256, 512 for
 * the winner.  Stage 3: waves per SIMD (no cap, 6, 5, 4), timed on one
 * batch.  Every candidate gives the same per-sample results.  Samples of one
 * size class make a representative input: a batch's best shape follows its
 * sample sizes.  results (may be NULL) receives up to max_results timed
 * candidates. */
int plb_batch_autotune(const plb_config *base, const plb_sample *samples, int n_samples, int steps,
                       int persistent, plb_config *best, plb_tune_result *results, int max_results,
                       int *n_results);

#ifdef __cplusplus
}
#endif
#endif
