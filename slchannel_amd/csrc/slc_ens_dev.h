// slc_ens_dev.h -- device-side work distribution of the batched ensemble
// kernels: contiguous XCD-aware chunks of the concatenated work units.
#pragma once

#include "slc_ens.h"

namespace slc {

__device__ inline void ens_span(unsigned total, unsigned& beg, unsigned& end) {
    const unsigned G = gridDim.x;
    const unsigned lb = (unsigned)xcd_block(G);
    const unsigned per = total / G, rem = total % G;
    beg = lb * per + (lb < rem ? lb : rem);
    end = beg + per + (lb < rem ? 1u : 0u);
}

// largest m with start[m] <= u
__device__ inline int ens_find(const unsigned* start, int nmem, unsigned u) {
    int lo = 0, hi = nmem - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= u) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// f(member index, unit index within the member) for every unit of this
// workgroup's chunk; a barrier separates consecutive units (LDS reuse).
template <typename Fn>
__device__ inline void ens_for(const unsigned* __restrict__ start, int nmem, unsigned total,
                               Fn&& f) {
    unsigned beg, end;
    ens_span(total, beg, end);
    if (beg >= end) return;
    int m = ens_find(start, nmem, beg);
    for (unsigned u = beg; u < end; ++u) {
        while (u >= start[m + 1]) ++m;
        if (u != beg) __syncthreads();
        f(m, u - start[m]);
    }
}

// Kernel parameters of a batched launch.  The member table is read through
// a restrict-qualified read-only kernel argument so its loads are invariant:
// the compiler may re-load member fields (scalar loads) instead of keeping
// them live across the tile body.
#define ENS_PARAMS \
    const EnsMember* __restrict__ mem, const unsigned* __restrict__ start, int nmem, \
        unsigned total
#define ENS_ARGS(L) (L).mem, (L).start, (L).nmem, (L).total



}  // namespace slc
