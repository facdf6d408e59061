// slc_step.hip -- LDS-tiled semi-Lagrangian step kernels for gfx950 (one
// channel per launch); the tile itself is sl_tile_body in slc_tile.h.
#include <atomic>

#include "slc_tile.h"

namespace slc {

// tile shape of the step kernels: 4 x 4 x 16 arrival points (one per
// thread), 24 KB of LDS, at least 4 waves per SIMD (measured at 256^3 order 6
// against 4x4x32 / 2x8x16 / 8x4x8 / 4x8x8 tiles and 20-40 KB budgets).
// Compile-time overridable with -DSLC_TILE_TI=... for tile-shape sweeps.
#ifndef SLC_TILE_TI
#define SLC_TILE_TI 4
#define SLC_TILE_TJ 4
#define SLC_TILE_TK 16
#define SLC_TILE_LDS 24576
#define SLC_TILE_WAVES 4
#endif

#ifndef SLC_TILE_WAVES
#define SLC_TILE_WAVES 4
#endif
// waves per SIMD the near pass and the gather at given feet are built for
// (order 6).  At order 4 they are built for seven, with a 21 KB box budget so
// that seven workgroups fit a CU's 160 KB of LDS (72 VGPRs, 16 B spilled per
// lane; measured at 256^3: six waves 0.446 ms, seven 0.415 ms, eight with a
// 19 KB budget spill 40 B per lane and take 0.445 ms).  The order-4 box of a
// 4x4x16 tile at the usual CFL is about 10 KB.
#ifndef SLC_TILE_WAVES_GATHER
#define SLC_TILE_WAVES_GATHER SLC_TILE_WAVES
#endif
#ifndef SLC_TILE_WAVES_GATHER4
#define SLC_TILE_WAVES_GATHER4 7
#endif
#ifndef SLC_GTILE_LDS4
#define SLC_GTILE_LDS4 21504
#endif
// waves per SIMD the fused far + near pass is built for
#ifndef SLC_TILE_WAVES_BDF2
#define SLC_TILE_WAVES_BDF2 SLC_TILE_WAVES
#endif
// tile of the near pass and the gather at given feet (default: the same)
#ifndef SLC_GTILE_TI
#define SLC_GTILE_TI SLC_TILE_TI
#define SLC_GTILE_TJ SLC_TILE_TJ
#define SLC_GTILE_TK SLC_TILE_TK
#define SLC_GTILE_LDS SLC_TILE_LDS
#endif

template <typename T, int C, int P, int MODE, int TI, int TJ, int TK, int LDS_BYTES, int WAVES>
__global__ __launch_bounds__(kSB) __attribute__((amdgpu_waves_per_eu(WAVES))) void sl_tile_kernel(
    Grid g, ArrivalTabs A, TrajVel V, const double* __restrict__ F,
    const double* __restrict__ Fm, const T* __restrict__ Sn, const T* __restrict__ Sm,
    double dt, double* __restrict__ out, Feet ft, unsigned long long* nclamp,
    unsigned* defer) {
    __shared__ __align__(16) T lds[LDS_BYTES / sizeof(T)];  // 16-byte pairs land here (stage_box)
    __shared__ int sh[MODE == MODE_BDF2 ? 12 : 6];
    __shared__ int s_cnt;
    __shared__ FootSt nst[MODE == MODE_BDF2 ? TI * TJ * TK : 1];  // near stencils (MODE_BDF2)
    sl_tile_body<T, C, P, MODE, TI, TJ, TK, LDS_BYTES>(g, A, V, F, Fm, Sn, Sm, dt, out, ft,
                                                       nclamp, xcd_block(gridDim.x), lds, sh,
                                                       &s_cnt, nst, defer);
}

// The gather at given feet for the tiles a first pass deferred (their boxes
// exceed its LDS budget; flags[tile] = 1).  Its own tiles are 8 x 8 x 4: for
// feet spread far from their arrival points in x and y (the +-5-cell random
// displacement of the 1024x1024x512 benchmark) a flat tile's box is smaller
// per point than a 4 x 4 x 16 tile's (23 x 23 x 9 against 20 x 20 x 21
// doubles at order 6), so a 40 KB budget holds it and four workgroups share a
// CU.  A wide tile runs when a first-pass tile it overlaps is flagged, and
// gathers only the points of flagged tiles (each point is gathered and its
// clamp counted once over the two passes).  A box that does not fit this
// budget either takes the direct-load path.  Measured at 512^3, order 6,
// +-5 cells: 4 x 4 x 16 wide tiles with 76 KB (two per CU) 16.9 ms, these
// 11.7 ms with one workgroup per tile and 12.6 ms persistent (below); every
// tile 8 x 8 x 4 in the first pass 10.3 ms, but the normal gather 16% slower.
constexpr int kWideTI = 8, kWideTJ = 8, kWideTK = 4;
constexpr int kWideLds = 40832;  // 160 KB / 4, less the block's other LDS
#ifndef SLC_WIDE_WAVES
#define SLC_WIDE_WAVES 4
#endif
constexpr int kWideBlocksPerCU = SLC_WIDE_WAVES;

// Whether wide tile wt overlaps a flagged first-pass tile (FI x FJ x FK
// tiles, numbered like the wide ones, z fastest).  Independent loads; indices
// past the domain's last tile repeat that tile.
template <int C>
__device__ inline bool wide_wanted(const Grid& g, const unsigned* __restrict__ flags, long wt) {
    constexpr int FI = SLC_GTILE_TI, FJ = SLC_GTILE_TJ, FK = SLC_GTILE_TK;
    static_assert(kWideTI % FI == 0 && kWideTJ % FJ == 0 && FK % kWideTK == 0,
                  "a wide tile covers whole first-pass tiles in x, y and lies in one in z");
    constexpr int NI = kWideTI / FI, NJ = kWideTJ / FJ;
    const int nzi = Comp<C>::NZI(g);
    const int nwk = (nzi + kWideTK - 1) / kWideTK, nwj = (g.ny + kWideTJ - 1) / kWideTJ;
    const int wk = (int)(wt % nwk);
    const long r = wt / nwk;
    const int wj = (int)(r % nwj), wi = (int)(r / nwj);
    const int i0 = wi * kWideTI, j0 = wj * kWideTJ;
    const int i1 = min(i0 + kWideTI, g.nx) - 1, j1 = min(j0 + kWideTJ, g.ny) - 1;
    const int ftk = (nzi + FK - 1) / FK, ftj = (g.ny + FJ - 1) / FJ;
    unsigned anyf = 0u;
#pragma unroll
    for (int a = 0; a < NI; ++a)
#pragma unroll
        for (int b = 0; b < NJ; ++b) {
            const int fi = min(i0 / FI + a, i1 / FI), fj = min(j0 / FJ + b, j1 / FJ);
            anyf |= flags[((long)fi * ftj + fj) * ftk + (wk * kWideTK) / FK];
        }
    return anyf != 0u;
}

// Persistent: a few workgroups per CU walk the wide tiles in chunks of kSB
// (each lane tests one tile, one round of loads per chunk) and gather the
// wanted tiles of a chunk one after the other, so that a gather with
// nothing deferred costs a few hundred short workgroups instead of one per
// tile: at 256^3 the one-per-tile launch spent 60-100 us dispatching
// workgroups that exit (order-6 gather 0.84 -> 0.79 ms, order 4 0.42 ->
// 0.36 ms with this one).  The loop keeps the kernel arguments live, so at
// order 6 the SGPRs overflow into VGPR lanes and 96 B per lane spill (+-5
// cells at 512^3: 12.6 ms against 11.7 ms one workgroup per tile; three
// workgroups per CU, without the spill, 14.2 ms).
template <int C, int P>
__global__ __launch_bounds__(kSB) __attribute__((amdgpu_waves_per_eu(SLC_WIDE_WAVES))) void gather_wide_kernel(
    Grid g, ArrivalTabs A, const double* __restrict__ F, Feet ft, double* __restrict__ out,
    unsigned long long* nclamp, const unsigned* __restrict__ flags, long wtiles) {
    __shared__ __align__(16) double lds[kWideLds / sizeof(double)];
    __shared__ int sh[6];
    __shared__ int s_cnt;
    __shared__ unsigned long long s_want[kSB / 64];
    const TrajVel V{{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    Redo redo;
    redo.flags = flags;
    redo.ti = SLC_GTILE_TI;
    redo.tj = SLC_GTILE_TJ;
    redo.tk = SLC_GTILE_TK;
    const long nchunks = (wtiles + kSB - 1) / kSB;
    for (long ch = xcd_block(gridDim.x); ch < nchunks; ch += gridDim.x) {
        const long wt = ch * kSB + threadIdx.x;
        const bool want = wt < wtiles && wide_wanted<C>(g, flags, wt);
        const unsigned long long m = __ballot(want);
        if ((threadIdx.x & 63) == 0) s_want[threadIdx.x >> 6] = m;
        __syncthreads();
        for (int w = 0; w < kSB / 64; ++w) {
            // uniform over the block: scalar loop control around the barriers
            // of the tile body
            const unsigned long long mw = s_want[w];
            unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)mw);
            unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(mw >> 32));
            while (lo | hi) {
                int bit;
                if (lo) {
                    bit = __builtin_ctz(lo);
                    lo &= lo - 1;
                } else {
                    bit = 32 + __builtin_ctz(hi);
                    hi &= hi - 1;
                }
                sl_tile_body<double, C, P, MODE_GATHER, kWideTI, kWideTJ, kWideTK, kWideLds>(
                    g, A, V, F, F, F, F, 0.0, out, ft, nclamp, ch * kSB + w * 64 + bit, lds, sh,
                    &s_cnt, nullptr, nullptr, redo);
            }
        }
        __syncthreads();  // s_want is rewritten by the next chunk
    }
}

// Compute units of the current device (cached per device).
static int device_cus() {
    static std::atomic<int> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int n = cache[dev].load(std::memory_order_relaxed);
    if (n > 0) return n;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        n <= 0)
        n = 256;
    cache[dev].store(n, std::memory_order_relaxed);
    return n;
}

// Arrival tables of component C: x/dx, y/dy of the arrival columns and the
// face cell and centre split of each arrival z (2 c + split), evaluated with
// the very functions the trajectory sampling uses (so looked-up values are
// the computed ones).
template <int C>
__global__ void arrival_tabs_kernel(Grid g, double* xdx, double* ydy, int* cell) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nzi = Comp<C>::NZI(g);
    double xa, ya, za;
    int I, J, K;
    if (t < g.nx) {
        arrival<C>(g, (long)t * g.ny * nzi, xa, ya, za, I, J, K);
        xdx[t] = xa / g.dx;
    }
    if (t < g.ny) {
        arrival<C>(g, (long)t * nzi, xa, ya, za, I, J, K);
        ydy[t] = ya / g.dy;
    }
    if (t < nzi) {
        arrival<C>(g, (long)t, xa, ya, za, I, J, K);
        const double zc = clampd(za, g.z_lo, g.z_hi);
        const int c = face_cell(g, zc);
        cell[t] = 2 * c + (zc >= g.z_c[c + 1] ? 1 : 0);
    }
}

hipError_t run_arrival_tabs(const Grid& g, int comp, double* xdx, double* ydy, int* cell,
                            hipStream_t s) {
    int n = g.nx > g.ny ? g.nx : g.ny;
    n = n > g.nz ? n : g.nz;
    dim3 grid((n + 255) / 256);
    if (comp == 0) hipLaunchKernelGGL(arrival_tabs_kernel<0>, grid, dim3(256), 0, s, g, xdx, ydy, cell);
    else if (comp == 1) hipLaunchKernelGGL(arrival_tabs_kernel<1>, grid, dim3(256), 0, s, g, xdx, ydy, cell);
    else hipLaunchKernelGGL(arrival_tabs_kernel<2>, grid, dim3(256), 0, s, g, xdx, ydy, cell);
    return hipGetLastError();
}

// U* = 2a - b elementwise over three whole ghost-shaped fields
// (solver.py:866-868: the same single rounding).
// With fa/fb non-null also writes float copies of a and b (FP32 storage).
struct F32x3 {
    float* p[3];
};

// U* = 2 V^n - V^{n-1} of the three components, one element per thread
// (a two-element, 16-byte variant measured 7% slower at 256^3).
__global__ void ustar_kernel(double* __restrict__ o0, double* __restrict__ o1,
                             double* __restrict__ o2, const double* __restrict__ a0,
                             const double* __restrict__ a1, const double* __restrict__ a2,
                             const double* __restrict__ b0, const double* __restrict__ b1,
                             const double* __restrict__ b2, long n0, long n1, long n2,
                             F32x3 fa, F32x3 fb) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool f32 = fa.p[0] != nullptr;
    if (i < n0) {
        const double a = a0[i], b = b0[i];
        o0[i] = 2.0 * a - b;
        if (f32) { fa.p[0][i] = (float)a; fb.p[0][i] = (float)b; }
    }
    if (i < n1) {
        const double a = a1[i], b = b1[i];
        o1[i] = 2.0 * a - b;
        if (f32) { fa.p[1][i] = (float)a; fb.p[1][i] = (float)b; }
    }
    if (i < n2) {
        const double a = a2[i], b = b2[i];
        o2[i] = 2.0 * a - b;
        if (f32) { fa.p[2][i] = (float)a; fb.p[2][i] = (float)b; }
    }
}

static inline void field_sizes(const Grid& g, long n[3]) {
    n[0] = (long)(g.nx + 1) * (g.ny + 2) * (g.nz + 2);
    n[1] = (long)(g.nx + 2) * (g.ny + 1) * (g.nz + 2);
    n[2] = (long)(g.nx + 2) * (g.ny + 2) * (g.nz + 1);
}

hipError_t run_ustar(const Grid& g, double* const o[3], const double* const a[3],
                     const double* const b[3], float* const fa[3], float* const fb[3],
                     hipStream_t s) {
    long n[3];
    field_sizes(g, n);
    long nmax = n[0] > n[1] ? n[0] : n[1];
    nmax = nmax > n[2] ? nmax : n[2];
    F32x3 A{{fa ? fa[0] : nullptr, fa ? fa[1] : nullptr, fa ? fa[2] : nullptr}};
    F32x3 B{{fb ? fb[0] : nullptr, fb ? fb[1] : nullptr, fb ? fb[2] : nullptr}};
    hipLaunchKernelGGL(ustar_kernel, dim3((unsigned)((nmax + 255) / 256)), dim3(256), 0, s, o[0],
                       o[1], o[2], a[0], a[1], a[2], b[0], b[1], b[2], n[0], n[1], n[2], A, B);
    return hipGetLastError();
}

// float copies of three ghost-shaped fields (FP32 storage for advect / BDF1)
__global__ void to_f32_kernel(float* __restrict__ o0, float* __restrict__ o1,
                              float* __restrict__ o2, const double* __restrict__ a0,
                              const double* __restrict__ a1, const double* __restrict__ a2,
                              long n0, long n1, long n2) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n0) o0[i] = (float)a0[i];
    if (i < n1) o1[i] = (float)a1[i];
    if (i < n2) o2[i] = (float)a2[i];
}

hipError_t run_to_f32(const Grid& g, float* const o[3], const double* const a[3],
                      hipStream_t s) {
    long n[3];
    field_sizes(g, n);
    long nmax = n[0] > n[1] ? n[0] : n[1];
    nmax = nmax > n[2] ? nmax : n[2];
    hipLaunchKernelGGL(to_f32_kernel, dim3((unsigned)((nmax + 255) / 256)), dim3(256), 0, s, o[0],
                       o[1], o[2], a[0], a[1], a[2], n[0], n[1], n[2]);
    return hipGetLastError();
}

constexpr int kTI = SLC_TILE_TI, kTJ = SLC_TILE_TJ, kTK = SLC_TILE_TK, kLDS = SLC_TILE_LDS;

struct StepArgs {
    const double* F;
    const double* Fm;
    const void* Sn;  // gathered copies (double or float)
    const void* Sm;
    double dt;
    double* out;
    Feet ft;
    unsigned long long* ncl;
    unsigned* defer;  // MODE_GATHER: per-tile flags of the wide second pass, or null
};

template <typename T, int C, int P, int MODE>
static void launch_mode(const Grid& g, const ArrivalTabs& A, const TrajVel& V,
                        const StepArgs& a, hipStream_t s) {
    constexpr bool G = MODE == MODE_NEAR || MODE == MODE_GATHER;
    constexpr int TI = G ? SLC_GTILE_TI : kTI, TJ = G ? SLC_GTILE_TJ : kTJ;
    constexpr int TK = G ? SLC_GTILE_TK : kTK;
    constexpr int LDS = G ? (P == 4 ? SLC_GTILE_LDS4 : SLC_GTILE_LDS) : kLDS;
    constexpr int W = G ? (P == 4 ? SLC_TILE_WAVES_GATHER4 : SLC_TILE_WAVES_GATHER)
                        : (MODE == MODE_BDF2 ? SLC_TILE_WAVES_BDF2 : SLC_TILE_WAVES);
    const long nzi = (C == 2) ? g.nz - 1 : g.nz;
    const unsigned tiles = (unsigned)(((long)(g.nx + TI - 1) / TI) *
                                      ((g.ny + TJ - 1) / TJ) * ((nzi + TK - 1) / TK));
    if (tiles == 0) return;
    unsigned* defer = (MODE == MODE_GATHER) ? a.defer : nullptr;
    hipLaunchKernelGGL((sl_tile_kernel<T, C, P, MODE, TI, TJ, TK, LDS, W>), dim3(tiles),
                       dim3(kSB), 0, s, g, A, V, a.F, a.Fm, (const T*)a.Sn, (const T*)a.Sm, a.dt,
                       a.out, a.ft, a.ncl, defer);
    if constexpr (MODE == MODE_GATHER && sizeof(T) == 8) {
        if (defer) {
            const long wtiles = ((long)(g.nx + kWideTI - 1) / kWideTI) *
                                ((g.ny + kWideTJ - 1) / kWideTJ) *
                                ((nzi + kWideTK - 1) / kWideTK);
            const long nchunks = (wtiles + kSB - 1) / kSB;
            const long cap = (long)device_cus() * kWideBlocksPerCU;
            hipLaunchKernelGGL((gather_wide_kernel<C, P>),
                               dim3((unsigned)(nchunks < cap ? nchunks : cap)), dim3(kSB), 0, s,
                               g, A, (const double*)a.Sn, a.ft, a.out, a.ncl, defer, wtiles);
        }
    }
}

template <typename T, int C, int MODE>
static void launch_cp(const Grid& g, const ArrivalTabs& A, const TrajVel& V, const StepArgs& a,
                      hipStream_t s) {
    if (g.order == 4) launch_mode<T, C, 4, MODE>(g, A, V, a, s);
    else launch_mode<T, C, 6, MODE>(g, A, V, a, s);
}

template <typename T, int MODE>
static void launch_c(const Grid& g, int comp, const ArrivalTabs& A, const TrajVel& V,
                     const StepArgs& a, hipStream_t s) {
    switch (comp) {
        case 0: launch_cp<T, 0, MODE>(g, A, V, a, s); break;
        case 1: launch_cp<T, 1, MODE>(g, A, V, a, s); break;
        default: launch_cp<T, 2, MODE>(g, A, V, a, s); break;
    }
}

template <typename T>
static void launch_t(const Grid& g, int comp, int mode, const ArrivalTabs& A, const TrajVel& V,
                     const StepArgs& a, hipStream_t s) {
    switch (mode) {
        case MODE_ADVECT: launch_c<T, MODE_ADVECT>(g, comp, A, V, a, s); break;
        case MODE_BDF1: launch_c<T, MODE_BDF1>(g, comp, A, V, a, s); break;
        case MODE_FAR: launch_c<T, MODE_FAR>(g, comp, A, V, a, s); break;
        case MODE_NEAR: launch_c<T, MODE_NEAR>(g, comp, A, V, a, s); break;
        case MODE_BDF2: launch_c<T, MODE_BDF2>(g, comp, A, V, a, s); break;
        default: launch_c<T, MODE_GATHER>(g, comp, A, V, a, s); break;
    }
}

hipError_t run_tile_step(const Grid& g, int comp, int mode, const ArrivalTabs& A,
                         const TrajVel& V, const double* F, const double* Fm, const float* Fn32,
                         const float* Fm32, double dt, double* out, double* fx, double* fy,
                         double* fz, unsigned long long* ncl, hipStream_t s, unsigned* defer) {
    const Feet ft{fx, fy, fz};
    if (Fn32 || Fm32) {
        StepArgs a{F, Fm, Fn32, Fm32, dt, out, ft, ncl, nullptr};
        launch_t<float>(g, comp, mode, A, V, a, s);
    } else {
        StepArgs a{F, Fm, F, Fm, dt, out, ft, ncl, defer};
        launch_t<double>(g, comp, mode, A, V, a, s);
    }
    return hipGetLastError();
}

long gather_tiles(const Grid& g, int comp) {
    const long nzi = (comp == 2) ? g.nz - 1 : g.nz;
    return ((long)(g.nx + SLC_GTILE_TI - 1) / SLC_GTILE_TI) *
           ((g.ny + SLC_GTILE_TJ - 1) / SLC_GTILE_TJ) * ((nzi + SLC_GTILE_TK - 1) / SLC_GTILE_TK);
}

}  // namespace slc
