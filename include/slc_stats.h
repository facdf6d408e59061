/* slc_stats.h -- on-device turbulence statistics of libslc_hip.so.
 *
 * Replaces the per-snapshot work of TurbulenceStats.accumulate_statistics
 * (slchannel/turbstats.py:156-361): plane means of the cell-centred
 * velocities and their time-mean-correction sums, u'u' and v'v' at the u and
 * v nodes, w'w' at the z faces averaged to centres, <u'w'> from cell-centred
 * fluctuations, and the folded 2-D spectra (uu, vv, ww, uw) at plane pairs
 * (bottom, top; w negated at the top, both walls averaged).  The caller owns
 * the running sums (device float64 buffers) and the sample counters;
 * finalize_statistics (turbstats.py:362-467) stays host arithmetic on nz
 * numbers.
 *
 * Plane reductions are two-pass (means, then moments about them) in a fixed
 * order, so the sums are deterministic run to run.  Fields are the ghost-
 * shaped float64 arrays of SLChannelFlow (u (nx+1, ny+2, nz+2), v (nx+2,
 * ny+1, nz+2), w (nx+2, ny+2, nz+1), z fastest).
 */
#ifndef SLC_STATS_H
#define SLC_STATS_H

#include "slc.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct slc_stats slc_stats;

/* Number of profile sums and their order in the prof buffer (each nz long):
 * U, uu, vv, ww, uw, Vm, Wm, Um2, Vm2, Wm2, UmWm -- the accumulators
 * U_sum .. UmWm_sum of turbstats.py:98-120. */
#define SLC_STATS_NPROF 11

/* Workspace for one grid.  nplanes spectra plane pairs with ghost-offset z
 * indices k_bot[i], k_top[i] in 1..nz (host arrays; the primary z+ plane and
 * the extra planes of turbstats.py:79-91, 142-154).  nplanes may be 0.
 * Allocates every buffer and rocFFT plan here; slc_stats_accumulate
 * allocates nothing.  Spectra need nx >= 2. */
int slc_stats_create(int device, int nx, int ny, int nz, int nplanes, const int* k_bot,
                     const int* k_top, slc_stats** out);
int slc_stats_destroy(slc_stats* st);
int slc_stats_workspace_bytes(const slc_stats* st, int64_t* bytes);

/* Adds one snapshot (turbstats.py:156-361) to the running sums:
 *   prof: device, SLC_STATS_NPROF x nz float64;
 *   spec: device, nplanes x 4 x (nx/2) x (ny/2) float64 (uu, vv, ww, uw),
 *         may be null when nplanes == 0.
 * Runs on `stream`; calls on one slc_stats must not overlap (they share its
 * workspace). */
int slc_stats_accumulate(slc_stats* st, const double* u, const double* v, const double* w,
                         double* prof, double* spec, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SLC_STATS_H */
