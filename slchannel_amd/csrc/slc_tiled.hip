// slc_tiled.hip -- LDS-tiled order-P gathers for gfx950.
//
// A workgroup owns a TI x TJ x TK tile of arrival points (z fastest).  Its
// feet are read once (coalesced), their stencil bases reduced to the tile's
// bounding box -- the halo is whatever the trajectory CFL makes it, measured
// per tile -- and the box of the source field is staged in LDS with
// coalesced, z-contiguous loads (periodic wrap in x/y).  Every point then
// contracts its P^3 stencil from LDS in the same order as the direct
// gather (z, then y, then x, with FMAs), so the two paths give identical
// results.  A tile whose box exceeds the LDS budget falls back to direct
// global loads for that tile only.
//
// Why: the direct gather issues P^3 scattered 8-byte loads per point through
// the vector L1 (64 B/clk/CU); from LDS the same reads cost half the cycles
// and the HBM/L2 side sees ~5-9 coalesced loads per point instead of 216.
#include <climits>

#include "slc_device.h"
#include "slc_internal.h"

namespace slc {

constexpr int kTB = 256;  // threads per workgroup

struct BBox {
    int xmin, xmax, ymin, ymax, kmin, kmax;
};

__device__ inline int wave_min(int v) {
    for (int o = 32; o > 0; o >>= 1) {
        int w = __shfl_xor(v, o);
        v = w < v ? w : v;
    }
    return v;
}

__device__ inline int wave_max(int v) {
    for (int o = 32; o > 0; o >>= 1) {
        int w = __shfl_xor(v, o);
        v = w > v ? w : v;
    }
    return v;
}

// Block-wide bounding box of the per-thread ranges (uniform result).
__device__ inline BBox block_bbox(BBox l, int* sh /* 6 ints of LDS */) {
    if (threadIdx.x == 0) {
        sh[0] = INT_MAX; sh[1] = INT_MIN;
        sh[2] = INT_MAX; sh[3] = INT_MIN;
        sh[4] = INT_MAX; sh[5] = INT_MIN;
    }
    __syncthreads();
    int a = wave_min(l.xmin), b = wave_max(l.xmax);
    int c = wave_min(l.ymin), d = wave_max(l.ymax);
    int e = wave_min(l.kmin), f = wave_max(l.kmax);
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&sh[0], a); atomicMax(&sh[1], b);
        atomicMin(&sh[2], c); atomicMax(&sh[3], d);
        atomicMin(&sh[4], e); atomicMax(&sh[5], f);
    }
    __syncthreads();
    BBox r{sh[0], sh[1], sh[2], sh[3], sh[4], sh[5]};
    __syncthreads();
    return r;
}

// Stage the box [xmin, xmin+BX) x [ymin, ymin+BY) x [kmin, kmin+BZ) of the
// component-C field F (ghost-shaped) into lds[(x*BY + y)*BZ + z].
template <int C>
__device__ inline void stage_box(const Grid& g, const double* __restrict__ F, double* lds,
                                 int xmin, int ymin, int kmin, int BX, int BY, int BZ) {
    const int rows = BX * BY;
    // one row (x, y) = BZ contiguous doubles; lanes stride along z
    for (int e = threadIdx.x; e < rows * BZ; e += kTB) {
        int row = e / BZ, z = e - row * BZ;
        int x = row / BY, y = row - x * BY;
        int gx = wrap_idx(xmin + x, g.nx), gy = wrap_idx(ymin + y, g.ny);
        lds[e] = F[node_addr<C>(g, gx, gy, kmin + z)];
    }
}

template <int C, int P, int TI, int TJ, int TK, int CAP>
__global__ __launch_bounds__(kTB) void gather_tiled_kernel(Grid g, const double* __restrict__ F,
                                                           const double* __restrict__ xd,
                                                           const double* __restrict__ yd,
                                                           const double* __restrict__ zd,
                                                           double* __restrict__ out,
                                                           unsigned long long* nclamp) {
    constexpr int NPT = TI * TJ * TK;
    static_assert(NPT % kTB == 0, "tile must be a multiple of the block");
    constexpr int PT = NPT / kTB;
    __shared__ double lds[CAP];
    __shared__ int sh[6];
    const int nzi = Comp<C>::NZI(g);
    const int ntk = (nzi + TK - 1) / TK, ntj = (g.ny + TJ - 1) / TJ;
    const long tile = xcd_block(gridDim.x);
    const int tk = (int)(tile % ntk);
    const long r0 = tile / ntk;
    const int tj = (int)(r0 % ntj), ti = (int)(r0 / ntj);
    const int i0 = ti * TI, j0 = tj * TJ, k0t = tk * TK;

    long idx[PT];
    int bx[PT], by[PT], bk[PT];
    BBox l{INT_MAX, INT_MIN, INT_MAX, INT_MIN, INT_MAX, INT_MIN};
    int ncl = 0;
#pragma unroll
    for (int p = 0; p < PT; ++p) {
        const int q = threadIdx.x + kTB * p;
        const int kk = q % TK, t2 = q / TK;
        const int jj = t2 % TJ, ii = t2 / TJ;
        const int i = i0 + ii, j = j0 + jj, k = k0t + kk;
        idx[p] = -1;
        if (i < g.nx && j < g.ny && k < nzi) {
            const long n = ((long)i * g.ny + j) * nzi + k;  // < 2^31 by plan limits
            idx[p] = n;
            double z = zd[n];
            ncl += (z < g.z_lo || z > g.z_hi) ? 1 : 0;
            z = clampd(z, g.z_lo, g.z_hi);
            double tx, ty;
            stencil_base<C, P>(g, xd[n], yd[n], z, bx[p], by[p], bk[p], tx, ty);
            l.xmin = min(l.xmin, bx[p]); l.xmax = max(l.xmax, bx[p]);
            l.ymin = min(l.ymin, by[p]); l.ymax = max(l.ymax, by[p]);
            l.kmin = min(l.kmin, bk[p]); l.kmax = max(l.kmax, bk[p]);
        }
    }
    const BBox b = block_bbox(l, sh);
    // 64-bit extents, each bounded before the product (feet far outside the
    // domain or non-finite ones would overflow int extents)
    const long ex = (long)b.xmax - b.xmin + P, ey = (long)b.ymax - b.ymin + P,
               ez = (long)b.kmax - b.kmin + P;
    const bool any = b.xmin <= b.xmax;
    const bool fits = any && ex <= CAP && ey <= CAP && ez <= CAP && ex * ey * ez <= CAP;
    const int BX = fits ? (int)ex : 0, BY = fits ? (int)ey : 1, BZ = fits ? (int)ez : 1;
    if (fits) stage_box<C>(g, F, lds, b.xmin, b.ymin, b.kmin, BX, BY, BZ);
    __syncthreads();

    const double* zn = znodes<C>(g);
    const double* dinv = Comp<C>::centers ? g.dinv_cP : g.dinv_fP;
#pragma unroll 1
    for (int p = 0; p < PT; ++p) {
        if (idx[p] < 0) continue;
        const long n = idx[p];
        double z = clampd(zd[n], g.z_lo, g.z_hi);
        double tx, ty;
        int cbx, cby, cbk;
        stencil_base<C, P>(g, xd[n], yd[n], z, cbx, cby, cbk, tx, ty);
        double wx[P], wy[P], wz[P];
        uniform_weights<P>(tx, wx);
        uniform_weights<P>(ty, wy);
        z_weights<P>(zn, dinv, z, bk[p], wz);
        double acc = 0.0;
        if (fits) {
            const int ox = bx[p] - b.xmin, oy = by[p] - b.ymin, oz = bk[p] - b.kmin;
#pragma unroll
            for (int i = 0; i < P; ++i) {
                double ai = 0.0;
#pragma unroll
                for (int j = 0; j < P; ++j) {
                    const double* col = lds + ((ox + i) * BY + (oy + j)) * BZ + oz;
                    double sj = 0.0;
#pragma unroll
                    for (int k = 0; k < P; ++k) sj = fma(wz[k], col[k], sj);
                    ai = fma(wy[j], sj, ai);
                }
                acc = fma(wx[i], ai, acc);
            }
        } else {
            // rare: box larger than the LDS budget; rolled loop keeps the
            // register budget of the LDS path
#pragma unroll 1
            for (int i = 0; i < P; ++i) {
                const int gx = wrap_idx(bx[p] + i, g.nx);
                double ai = 0.0;
#pragma unroll
                for (int j = 0; j < P; ++j) {
                    const int gy = wrap_idx(by[p] + j, g.ny);
                    const double* col = F + node_addr<C>(g, gx, gy, bk[p]);
                    double sj = 0.0;
#pragma unroll
                    for (int k = 0; k < P; ++k) sj = fma(wz[k], col[k], sj);
                    ai = fma(wy[j], sj, ai);
                }
                acc = fma(uniform_weight_m<P>(tx, i), ai, acc);
            }
        }
        out[n] = acc;
    }
    if (nclamp) {
        __shared__ int s_cnt;
        if (threadIdx.x == 0) s_cnt = 0;
        __syncthreads();
        if (ncl) atomicAdd(&s_cnt, ncl);
        __syncthreads();
        if (threadIdx.x == 0 && s_cnt) atomicAdd(nclamp, (unsigned long long)s_cnt);
    }
}

// Tile shapes (points per workgroup = TI*TJ*TK) and LDS budgets (doubles).
template <int C, int P>
static void launch_tiled(const Grid& g, int shape, const double* F, const double* xd,
                         const double* yd, const double* zd, double* out,
                         unsigned long long* ncl, hipStream_t s) {
    auto tiles = [&](int TI, int TJ, int TK) {
        long nz_i = (C == 2) ? g.nz - 1 : g.nz;
        return (unsigned)(((long)(g.nx + TI - 1) / TI) * ((g.ny + TJ - 1) / TJ) *
                          ((nz_i + TK - 1) / TK));
    };
    switch (shape) {
        case 0:
            hipLaunchKernelGGL((gather_tiled_kernel<C, P, 4, 8, 32, 8192>), dim3(tiles(4, 8, 32)),
                               dim3(kTB), 0, s, g, F, xd, yd, zd, out, ncl);
            break;
        case 1:
            hipLaunchKernelGGL((gather_tiled_kernel<C, P, 2, 8, 32, 5120>), dim3(tiles(2, 8, 32)),
                               dim3(kTB), 0, s, g, F, xd, yd, zd, out, ncl);
            break;
        case 2:
            hipLaunchKernelGGL((gather_tiled_kernel<C, P, 2, 4, 64, 5120>), dim3(tiles(2, 4, 64)),
                               dim3(kTB), 0, s, g, F, xd, yd, zd, out, ncl);
            break;
        default:
            hipLaunchKernelGGL((gather_tiled_kernel<C, P, 4, 4, 32, 5120>), dim3(tiles(4, 4, 32)),
                               dim3(kTB), 0, s, g, F, xd, yd, zd, out, ncl);
            break;
    }
}

template <int C>
static void launch_tiled_c(const Grid& g, int shape, const double* F, const double* xd,
                           const double* yd, const double* zd, double* out,
                           unsigned long long* ncl, hipStream_t s) {
    if (g.order == 4) launch_tiled<C, 4>(g, shape, F, xd, yd, zd, out, ncl, s);
    else launch_tiled<C, 6>(g, shape, F, xd, yd, zd, out, ncl, s);
}

hipError_t run_gather_tiled(const Grid& g, int comp, int shape, const double* F,
                            const double* xd, const double* yd, const double* zd, double* out,
                            unsigned long long* ncl, hipStream_t s) {
    switch (comp) {
        case 0: launch_tiled_c<0>(g, shape, F, xd, yd, zd, out, ncl, s); break;
        case 1: launch_tiled_c<1>(g, shape, F, xd, yd, zd, out, ncl, s); break;
        default: launch_tiled_c<2>(g, shape, F, xd, yd, zd, out, ncl, s); break;
    }
    return hipGetLastError();
}

}  // namespace slc
