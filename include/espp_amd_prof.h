/*
 * espp_amd_prof.h -- instrumentation entry points of libespp_amd.so.
 *
 * These have no counterpart in the reference: they serve measurement
 * (benchmarks, roofline accounting) and are not needed by a drop-in user.
 * Same conventions as espp_amd.h (0 on success, espp_last_error on failure).
 */
#ifndef ESPP_AMD_PROF_H
#define ESPP_AMD_PROF_H

#include <stdint.h>

#include "espp_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* kernel ids for espp_prof_stats */
#define ESPP_K_PAIR 0        /* Lennard-Jones pair-force kernel */
#define ESPP_K_NBR_BUILD 1   /* Verlet-list build */
#define ESPP_K_INTEGRATE1 2  /* thermostat + second half kick of the previous step + half
                                kick + drift + max displacement, where the previous step's
                                force kernel did not do it */
#define ESPP_K_INTEGRATE2 3  /* thermostat + half kick after the last step of a run */
#define ESPP_K_BINNING 4     /* fold, counting sort and permutation (all kernels of one binning) */
#define ESPP_K_BOND 5        /* FENE bond-force kernel */
#define ESPP_K_PAIR_STEP 6   /* Lennard-Jones pair forces that also finish the step: thermostat,
                                half kick, and the next step's half kick, drift and max
                                displacement (every step of a run but the last), walking
                                the full list and writing its pruned copy */
#define ESPP_K_PAIR_STEP_PRUNED 7 /* the same walking the pruned list */

/* hipDeviceSynchronize on the current device. */
int espp_device_synchronize(void);

/* Record HIP events around every launch of the kernels above on the handle's
 * stream (on != 0) or stop recording (on == 0). */
int espp_prof_enable(espp_sim *sim, int32_t on);

/* Clear the accumulated launch counts and times. */
int espp_prof_reset(espp_sim *sim);

/* Launch count and summed event time (ms) of one kernel id since the last reset. */
int espp_prof_stats(espp_sim *sim, int32_t kernel, int64_t *launches, double *total_ms);

/* Total neighbour entries of the full list (each pair counted from both ends),
 * the padded row capacity and the largest row of the last build. */
int espp_vl_stats(espp_sim *sim, int32_t vl, int64_t *entries, int32_t *cap, int32_t *maxcount);

/* Batches of `batch` list entries in the list's lane order (lanes of a tile
 * ordered as the pair kernels take them): summed over lanes, and summed over
 * wavefronts of 64 consecutive lanes with each wavefront counted at its longest
 * lane, which is what a wavefront runs. */
int espp_vl_batch_stats(espp_sim *sim, int32_t vl, int32_t batch, int64_t *lane_batches, int64_t *wave_batches);

/* Longest window of steps that espp_run launches without waiting on the host
 * (default 32; 0 or 1: a host wait after every step).  Within a window each
 * step's kernels first read a device flag, raised by the previous step's
 * displacement reduction once the summed displacement bound (the skin check's
 * maxDist) passes skin/2, and return at once when it is set; the host reads how
 * many steps ran, rebuilds and goes on.  Results are those of the step-by-step
 * loop.  Used when every force of a step is a tiled mixed-precision Lennard-Jones
 * or a FENE interaction, pruned lists are off and launch timing is off. */
int espp_set_step_window(espp_sim *sim, int32_t max_steps);

/* Windows run, steps launched in them and steps that ran, since the handle was made. */
int espp_step_window_stats(espp_sim *sim, int64_t *windows, int64_t *launched, int64_t *ran);

/* Width w of the shell that the pruned Lennard-Jones lists keep beyond the
 * largest cutoff (0 or at least the skin: no pruning, the default; below 0:
 * skin / 3).  A force walk of the mixed path over the full list also writes
 * the pairs within rc + w at its positions as the pruned list; the following
 * walks take the pruned list until twice the displacement bound accumulated
 * since (the skin check's maxDist) reaches w.  It pays only where lists live
 * many steps: the pruning walk runs at lower occupancy than the plain one. */
int espp_set_prune_shell(espp_sim *sim, double width);

/* Step kernels that pruned and that walked a pruned list (all interactions,
 * since the handle was made), and the entries of interaction ia's current
 * pruned list (0 when it has none). */
int espp_prune_stats(espp_sim *sim, int32_t ia, int64_t *prunes, int64_t *pruned_walks, int64_t *entries);

/* Particles per tile of the current tile plan (0: untiled kernels) and the
 * largest staged region in particles. */
int espp_tile_stats(espp_sim *sim, int32_t *tile, int32_t *max_region);

#ifdef __cplusplus
}
#endif

#endif /* ESPP_AMD_PROF_H */
