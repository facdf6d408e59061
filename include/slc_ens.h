/* slc_ens.h -- batched ensemble / serving layer of libslc_hip.so.
 *
 * An ensemble holds many independent channel runs ("members"), each of its
 * own grid size, Reynolds number and time step, resident in HBM, and
 * advances all of them by one SL-BDF2 step per slc_ens_step call with ONE
 * batched launch per substep (ghost fills, U*, far and near SL passes per
 * component, z-solves, divergence, Poisson Thomas, projection, bulk shift)
 * plus one batched rocFFT execution per distinct grid size.  Members can be
 * added and removed between steps (continuous batching); a newly added
 * member takes the BDF1 bootstrap on its first step while the others take
 * BDF2 steps in the same launches.
 *
 * Each member's step is the single-run step of SLChannelFlow.step_sl_bdf2
 * (solver.py:813-1018; the same device bodies as slc_bdf2_predict,
 * slc_zsolve, slc_project_step and slc_bulk_force), so a member's result does
 * not depend on the other members.
 *
 * The reference has no counterpart: it steps one run per process
 * (solver.py:1280-1500 run loop).  This API serves callers that advance many
 * runs at once: ensembles of uneven members and a serving loop with
 * continuous batching.
 *
 * Ownership: the ensemble allocates and owns every member's fields; callers
 * copy in (slc_ens_set_fields) and out (slc_ens_get_fields) with device
 * pointers.  Stream use: every call orders its work after `stream` and makes
 * `stream` wait for it; the ensemble runs on an internal stream of default
 * priority.  Not thread-safe: one host thread per ensemble.
 */
#ifndef SLC_ENS_H
#define SLC_ENS_H

#include "slc.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct slc_ens slc_ens;

/* order: 4 or 6 for every member; storage_bits: 64, or 32 for FP32-storage
 * gathers (SLC_OPT_FIELD_STORAGE). */
int slc_ens_create(int device, int order, int storage_bits, slc_ens** out);
int slc_ens_destroy(slc_ens* ens);

/* Device bytes one member of this grid would hold (fields, histories,
 * tables, its share of the spectral buffers), for admission by HBM
 * footprint.  desc->z_f / z_c may be null here. */
int slc_ens_member_bytes(const slc_grid_desc* desc, int storage_bits, int64_t* bytes);

/* Adds a member with zero fields; its first step is a BDF1 bootstrap.
 * desc->order must equal the ensemble's order; desc->nu = 1/Re. */
int slc_ens_add(slc_ens* ens, const slc_grid_desc* desc, double dt, double U_bulk, int* id);
int slc_ens_remove(slc_ens* ens, int id);
int slc_ens_count(const slc_ens* ens, int* n);
/* Total device bytes held by the ensemble (members, class buffers, FFT work
 * area, and the released member arenas kept for reuse). */
int slc_ens_bytes(const slc_ens* ens, int64_t* bytes);
/* Frees released member arenas, oldest first, until the pool holds at most
 * keep_bytes; *pool_bytes (optional) receives what the pool still holds.
 * keep_bytes < 0 only reports.  Freeing waits for the device (hipFree). */
int slc_ens_trim_pool(slc_ens* ens, int64_t keep_bytes, int64_t* pool_bytes);

/* Copies ghost-shaped u, v, w, p (device pointers, float64, SLChannelFlow
 * shapes) into the member and resets it to a BDF1 bootstrap (a restart,
 * solver.py:834-844).  p may be null (zero pressure). */
int slc_ens_set_fields(slc_ens* ens, int id, const double* u, const double* v, const double* w,
                       const double* p, void* stream);
/* Copies the member's current u, v, w, p out (device pointers). */
int slc_ens_get_fields(slc_ens* ens, int id, double* u, double* v, double* w, double* p,
                       void* stream);
/* Copies the member's statistics of its last step into out (device pointer,
 * 5 doubles): u_bulk after the shift, forcing, clamped feet, steps done,
 * u_bulk before the shift. */
int slc_ens_stats(slc_ens* ens, int id, double* out, void* stream);

/* Run-loop diagnostics of the member's current fields (slc_diagnostics:
 * max|div|, u_tau, CFL at the member's dt, dti, non-finite entries) into out
 * (5 device doubles), ordered on `stream` like the other calls: the serving
 * loop's NaN/blow-up guard without a host synchronisation. */
int slc_ens_diagnostics(slc_ens* ens, int id, double* out, void* stream);

/* One SL-BDF2 step of every member. */
int slc_ens_step(slc_ens* ens, void* stream);

/* Options:
 * SLC_ENS_OPT_TILE       SL tile shape 0: 4x4x16 (24 KB LDS), 1: 4x4x32 (40 KB),
 *                        2: 2x8x16 (24 KB)
 * SLC_ENS_OPT_GRID       workgroups of the non-SL kernels: 0 = one per work unit
 *                        (default), k = 1..32 persistent workgroups per CU
 * SLC_ENS_OPT_GRAPH      1: replay each step from a captured hipGraph (re-captured
 *                        after membership changes), 0: launch eagerly
 * SLC_ENS_OPT_BDF2_FUSED 1 (default): one SL tile pass per component (both feet
 *                        and gathers of the BDF2 members), 0: far pass, then
 *                        near pass (as SLC_OPT_BDF2_FUSED; same doubles)
 * SLC_ENS_OPT_POOL       released member arenas kept for reuse by a later
 *                        member of the same geometry (default 8, >= 0); beyond
 *                        it slc_ens_remove frees the oldest, and hipFree waits
 *                        for the device.  slc_ens_bytes counts the pool;
 *                        slc_ens_trim_pool frees it on demand.  (No counterpart
 *                        in the reference, which has no ensemble.) */
#define SLC_ENS_OPT_TILE 1
#define SLC_ENS_OPT_GRID 2
#define SLC_ENS_OPT_GRAPH 3
#define SLC_ENS_OPT_BDF2_FUSED 4
#define SLC_ENS_OPT_POOL 5
int slc_ens_set_option(slc_ens* ens, int option, int value);
int slc_ens_get_option(const slc_ens* ens, int option, int* value);

/* Autotuner: times the SL passes (far + near, three components) of the
 * current members `reps` times for every tile shape, then the other batched
 * substeps that write only scratch for grid modes 0, 4 and 16, and keeps the
 * fastest of each.  Leaves every member's state unchanged (the timed
 * kernels only write the predictor scratch, U*, the feet and spectral
 * buffers; clamp counters are cleared).  report (optional): "shape ms\n"
 * and "grid<k> ms\n" lines, each time printed exactly (%.9g of the float
 * that was compared; equal times keep the first candidate). */
int slc_ens_autotune(slc_ens* ens, int reps, void* stream, char* report, int64_t len);

#ifdef __cplusplus
}
#endif

#endif /* SLC_ENS_H */
