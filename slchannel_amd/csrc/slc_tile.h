// slc_tile.h -- the LDS-tiled "trace, then gather" tile of the semi-Lagrangian
// step, shared by the single-run kernels (slc_step.hip) and the batched
// ensemble kernels (slc_ens_kernels.hip).
//
// One kernel template covers the four uses of "trace, then gather":
//   MODE_ADVECT  foot at dt of the given trajectory velocity, gather F
//                (SLAdvector.advect, semilag.py:584-671)
//   MODE_BDF1    as ADVECT with U* = V^n, plus dt * nu L_xy V^n
//                (solver.py:737-765)
//   MODE_FAR     BOTH feet of the frozen U* (depths dt and 2dt, each an
//                independent iterated midpoint from the arrival point; their
//                first sample, at the arrival point itself, is shared), the
//                near foot stored for MODE_NEAR, V^{n-1} gathered at the far
//                foot into out                          (solver.py:870-882)
//   MODE_NEAR    V^n gathered at the stored near foot, then
//                out = 4/3 near - 1/3 out + (2dt/3)(2 R^n - R^{n-1})
//                                                       (solver.py:885-903)
//   MODE_BDF2    MODE_FAR and MODE_NEAR in one pass: the near foot's stencil
//                waits in LDS while the far box is staged and contracted,
//                then the near box is staged into the same buffer; the far
//                value stays in a register.  The same operations as the two
//                passes, so the same doubles, without the near foot and the
//                far value going through HBM.
// A workgroup owns a TI x TJ x TK tile of arrival points (z fastest).  Phase 1
// traces the tile's feet and keeps each foot's stencil (base indices,
// fractional offsets, clamped z) in registers; the tile's stencil bounding
// box (its trajectory-CFL halo, measured) is reduced and staged in LDS with
// coalesced z-contiguous loads; phase 2 contracts every P^3 stencil from LDS
// (z, then y, then x, with FMAs -- the same order as the direct gather).  A
// box larger than the LDS budget falls back to global loads for that tile.
#pragma once

#include <climits>

#include "slc_device.h"
#include "slc_internal.h"

namespace slc {

// threads per workgroup (compile-time overridable for tile sweeps)
#ifndef SLC_TILE_NT
#define SLC_TILE_NT 256
#endif
constexpr int kSB = SLC_TILE_NT;

enum { MODE_ADVECT = 0, MODE_BDF1 = 1, MODE_FAR = 2, MODE_NEAR = 3, MODE_GATHER = 4,
       MODE_BDF2 = 5 };

// Foot coordinates of one component (flat arrival order): written by
// MODE_FAR (near feet), read by MODE_NEAR and MODE_GATHER.
struct Feet {
    double* x;
    double* y;
    double* z;
};

struct SBox {
    int xmin, xmax, ymin, ymax, kmin, kmax;
};

// Wavefront min / max with DPP lane moves (registers only): swaps within
// quads, half-row and row mirrors give every lane its 16-lane row's value;
// row broadcasts of lanes 15 and 31 fold the rows into lane 63, read back as
// a uniform value.  A lane with no DPP source receives the operation's
// identity (INT_MAX for min, INT_MIN for max), which lets the compiler fold
// each lane move into the min / max itself (one DPP-modified instruction).
template <bool MIN>
__device__ inline int s_wave_reduce(int v) {
    constexpr int id = MIN ? INT_MAX : INT_MIN;
    auto op = [](int a, int b) { return MIN ? min(a, b) : max(a, b); };
    v = op(v, __builtin_amdgcn_update_dpp(id, v, 0xB1, 0xF, 0xF, false));   // quad_perm 1,0,3,2
    v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x4E, 0xF, 0xF, false));   // quad_perm 2,3,0,1
    v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x141, 0xF, 0xF, false));  // row_half_mirror
    v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x140, 0xF, 0xF, false));  // row_mirror
    v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x142, 0xA, 0xF, false));  // row_bcast:15
    v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x143, 0xC, 0xF, false));  // row_bcast:31
    return __builtin_amdgcn_readlane(v, 63);
}

__device__ inline int s_wave_min(int v) { return s_wave_reduce<true>(v); }

__device__ inline int s_wave_max(int v) { return s_wave_reduce<false>(v); }

// The workgroup unions of two boxes at once (one pair of barriers): b's left
// in sh[0..5] and returned (wave-uniform), bn's left in sh[6..11].
__device__ inline SBox s_block_box2(SBox l, SBox ln, int* sh) {
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 12; q += 2) {
            sh[q] = INT_MAX;
            sh[q + 1] = INT_MIN;
        }
    }
    __syncthreads();
    int v[12] = {s_wave_min(l.xmin), s_wave_max(l.xmax), s_wave_min(l.ymin),
                 s_wave_max(l.ymax), s_wave_min(l.kmin), s_wave_max(l.kmax),
                 s_wave_min(ln.xmin), s_wave_max(ln.xmax), s_wave_min(ln.ymin),
                 s_wave_max(ln.ymax), s_wave_min(ln.kmin), s_wave_max(ln.kmax)};
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int q = 0; q < 12; q += 2) {
            atomicMin(&sh[q], v[q]);
            atomicMax(&sh[q + 1], v[q + 1]);
        }
    }
    __syncthreads();
    return SBox{sh[0], sh[1], sh[2], sh[3], sh[4], sh[5]};
}

__device__ inline SBox s_block_box(SBox l, int* sh) {
    if (threadIdx.x == 0) {
        sh[0] = INT_MAX; sh[1] = INT_MIN;
        sh[2] = INT_MAX; sh[3] = INT_MIN;
        sh[4] = INT_MAX; sh[5] = INT_MIN;
    }
    __syncthreads();
    int a = s_wave_min(l.xmin), b = s_wave_max(l.xmax);
    int c = s_wave_min(l.ymin), d = s_wave_max(l.ymax);
    int e = s_wave_min(l.kmin), f = s_wave_max(l.kmax);
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&sh[0], a); atomicMax(&sh[1], b);
        atomicMin(&sh[2], c); atomicMax(&sh[3], d);
        atomicMin(&sh[4], e); atomicMax(&sh[5], f);
    }
    __syncthreads();
    SBox r{sh[0], sh[1], sh[2], sh[3], sh[4], sh[5]};
    __syncthreads();
    return r;
}

// Trajectory samples of the three U* components at one point (they share
// x/dx, y/dy and the face cell of z; semilag.py:512-528 recomputes them with
// identical results).  Two rounds of loads: the faces of the estimated
// interval, then the corners.
// xdx = x/dx, ydy = y/dy and zc the clamped z.
__device__ inline void sample3_exact(const Grid& g, const TrajA& V, double xdx, double ydy,
                                     double zc, double& us, double& vs, double& ws) {
    double split;
    const int c = face_cell_fast(g, zc, split);
    const int mc = locate_split<0>(c, split, zc);
    us = trilinear_m<0>(g, V, xdx, ydy, zc, mc);
    vs = trilinear_m<1>(g, V, xdx, ydy, zc, mc);
    ws = trilinear_m<2>(g, V, xdx, ydy, zc, c);
}

// Column values of component C at levels k0, k0+1 of its four trilinear
// columns, with the x/y weights of trilinear_m.
template <int C>
__device__ inline void trilinear_cols(const Grid& g, const TrajA& V, double xdx, double ydy,
                                      int k0, double* wx, double* wy, double (&v)[2][2][2]) {
    int ix[2], iy[2];
    trilinear_xy<C>(g, xdx, ydy, wx, wy, ix, iy);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int base = node_addr<C, false>(g, ix[i], iy[j], k0);
            v[i][j][0] = V.a[C][base];
            v[i][j][1] = V.a[C][base + 1];
        }
}

// trilinear_m's sum, over column values read by trilinear_cols.
__device__ inline double trilinear_sum(const double* wx, const double* wy, const double* wz,
                                       const double (&v)[2][2][2]) {
    double out = 0.0;
    bool first = true;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            double wij = wx[i] * wy[j];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                double term = (wij * wz[k]) * v[i][j][k];
                out = first ? term : out + term;
                first = false;
            }
        }
    }
    return out;
}

// sample3_exact in one round of loads, for a point traced from the arrival
// (za, K) of component C.  The face interval c and the centred components'
// node interval m = c + s (s: z >= z_c[c+1]) are predicted without loads, and
// the faces, centre nodes, inverse denominators and the three components'
// columns of that prediction are read at once (every index inside its table
// or field for c in [0, nz-1], s in {0, 1}).  Near the arrival the
// prediction is exact: for C = u, v, za = z_c[K] is the split of face
// interval K-1; for C = w, za = z_f[K] is the face between intervals K-1 and
// K.  Elsewhere c is the float estimate of face_cell_fast and s comes from
// split_predict.  The loaded faces and centre then decide: where they confirm
// the prediction and z is outside the faces' 1e-10 band, c is the interval
// face_cell_fast returns (the one interval bracketing z) and m the node
// interval of locate_split.  Otherwise (a few lanes in 1e5 on stretched
// production grids) c and m are located as sample3_exact does and the values
// read again.  Either way the sums are trilinear_m's operations on the same
// values, so the result is sample3_exact's.
template <int C>
__device__ inline void sample3(const Grid& g, const TrajA& V, double x, double y, double z,
                               double za, int K, double& us, double& vs, double& ws) {
    const double zc = clampd(z, g.z_lo, g.z_hi);
    const double xdx = x / g.dx, ydy = y / g.dy;
    float kfe, a;
    int c = face_estimate(g, zc, kfe, a);
    if (!Comp<C>::centers && (c == K - 1 || c == K)) c = zc >= za ? K : K - 1;
    bool s = split_predict(g, c, kfe, a);
    if (Comp<C>::centers && c == K - 1) s = zc >= za;
    int m = c + (s ? 1 : 0);
    for (bool located = false;; located = true) {
        const double lo = g.z_f[c], hi = g.z_f[c + 1];
        const double n0 = g.z_c[m], n1 = g.z_c[m + 1];
        const double p0 = g.dinv_c2[2 * m], p1 = g.dinv_c2[2 * m + 1];
        const double f0 = g.dinv_f2[2 * c], f1 = g.dinv_f2[2 * c + 1];
        double uwx[2], uwy[2], vwx[2], vwy[2], wwx[2], wwy[2];
        double uv[2][2][2], vv[2][2][2], wv[2][2][2];
        trilinear_cols<0>(g, V, xdx, ydy, m, uwx, uwy, uv);
        trilinear_cols<1>(g, V, xdx, ydy, m, vwx, vwy, vv);
        trilinear_cols<2>(g, V, xdx, ydy, c, wwx, wwy, wv);
        const double split = m > c ? n0 : n1;  // z_c[c + 1]
        if (located || (face_estimate_holds(g, c, zc, lo, hi) && (zc >= split) == (m > c))) {
            const double wzc[2] = {(zc - n1) * p0, (zc - n0) * p1};
            const double wzf[2] = {(zc - hi) * f0, (zc - lo) * f1};
            us = trilinear_sum(uwx, uwy, wzc, uv);
            vs = trilinear_sum(vwx, vwy, wzc, vv);
            ws = trilinear_sum(wwx, wwy, wzf, wv);
            return;
        }
        double sp;
        c = face_cell_fast(g, zc, sp);
        m = locate_split<0>(c, sp, zc);
    }
}

// Arrival-point trajectory sample with precomputed x/dx, y/dy and z location
// (csel = 2 * face cell + centre split: the values sample3_exact computes; see
// ArrivalTabs).
__device__ inline void sample3_at(const Grid& g, const TrajA& V, double xdx, double ydy,
                                  double z, int csel, double& us, double& vs, double& ws) {
    const double zc = clampd(z, g.z_lo, g.z_hi);
    const int c = csel >> 1, mc = c + (csel & 1);
    us = trilinear_m<0>(g, V, xdx, ydy, zc, mc);
    vs = trilinear_m<1>(g, V, xdx, ydy, zc, mc);
    ws = trilinear_m<2>(g, V, xdx, ydy, zc, c);
}

// Per-foot stencil data kept between the phases.
struct FootSt {
    int bx, by, bk;      // unwrapped x/y base, clamped z base
    double tx, ty, z;    // fractional offsets, clamped z
};

template <int C, int P>
__device__ inline void make_st(const Grid& g, double x, double y, double z, FootSt& f,
                               int& ncl) {
    ncl += (z < g.z_lo || z > g.z_hi) ? 1 : 0;
    f.z = clampd(z, g.z_lo, g.z_hi);
    stencil_base<C, P>(g, x, y, f.z, f.bx, f.by, f.bk, f.tx, f.ty);
}

// Stages the stencil bounding box `b` of a tile's feet (field S) into `lds`,
// using at most cap_bytes: the box (z fastest), then one int per (x, y) row,
// its global base.  Returns false (nothing staged) when the box does not fit.
// b.kmin becomes the z index of the box's first level.  Ends with a barrier.
//
// Double fields whose z rows have an even length and whose base is 16-byte
// aligned are staged with gfx950's direct global-to-LDS loads of 16-byte
// pairs (global_load_lds_dwordx4: no register round trip, no LDS write
// instruction): the box then starts at an even level and has an even
// z extent, so every pair lies inside one row (the rounded-up end is at most
// the row length, which is even) and is aligned.  The barrier's release
// fence waits for the loads.  Other fields (the w component at even nz, FP32
// storage) are staged through registers.
#ifndef SLC_STAGE_DMA
#define SLC_STAGE_DMA 1
#endif
template <typename T, int C, int P>
__device__ inline bool stage_box(const Grid& g, const T* __restrict__ S, SBox& b,
                                 T* __restrict__ lds, long cap_bytes, int& BY, int& BZ) {
    bool dma = false;
    if constexpr (SLC_STAGE_DMA && sizeof(T) == 8)
        dma = (Comp<C>::SZ(g) % 2 == 0) && ((reinterpret_cast<uintptr_t>(S) & 15) == 0);
    // extents in 64 bits, each bounded before the product: feet far outside
    // the domain (or non-finite ones, whose floors saturate) would overflow
    // int extents and must take the direct-load path
    const long ex = (long)b.xmax - b.xmin + P, ey = (long)b.ymax - b.ymin + P;
    auto fits_z = [&](long ez) {
        return b.xmin <= b.xmax && ex <= cap_bytes && ey <= cap_bytes && ez <= cap_bytes &&
               ex * ey * ez * (long)sizeof(T) + ex * ey * 4 <= cap_bytes;
    };
    // the even start and extent can add two levels: a box that fits only
    // without them is staged through registers
    if (dma) {
        const long ezd = (long)b.kmax - (b.kmin & ~1) + P;
        dma = fits_z(ezd + (ezd & 1));
    }
    const int k0 = dma ? (b.kmin & ~1) : b.kmin;
    long ez = (long)b.kmax - k0 + P;
    if (dma) ez += ez & 1;
    const bool fits = fits_z(ez);
    const int BX = fits ? (int)ex : 0;
    BY = fits ? (int)ey : 1;
    BZ = fits ? (int)ez : 1;
    const int rows = BX * BY;
    const int nbox = rows * BZ;
    if (fits) {
        b.kmin = k0;
        int* rowbase = reinterpret_cast<int*>(lds + nbox);
        for (int r = threadIdx.x; r < rows; r += kSB) {
            const int x = r / BY, y = r - (r / BY) * BY;
            rowbase[r] = node_addr<C>(g, wrap_idx(b.xmin + x, g.nx),
                                      wrap_idx(b.ymin + y, g.ny), k0);
        }
        __syncthreads();
        if (dma) {
            // pair q = 2 elements; a wavefront's 64 lanes load 64 consecutive
            // pairs into 1 KB of consecutive LDS per instruction; (row, zz) of
            // element 2q carried from one iteration to the next as below
            const int npairs = nbox / 2;
            const int lane = (int)threadIdx.x & 63;
            const int step = 2 * kSB;  // elements per iteration
            const int dq = step / BZ, dr = step - dq * BZ;
            int q = (int)threadIdx.x;
            int row = 2 * q / BZ;
            int zz = 2 * q - row * BZ;
#ifndef SLC_STAGE_DMA_UNROLL
#define SLC_STAGE_DMA_UNROLL 2  // 1, 2, 4 measured the same at 256^3
#endif
#pragma unroll SLC_STAGE_DMA_UNROLL
            for (int base = q - lane; base < npairs; base += kSB) {
                if (q < npairs)
                    __builtin_amdgcn_global_load_lds(
                        (const __attribute__((address_space(1))) void*)(S + rowbase[row] + zz),
                        (__attribute__((address_space(3))) void*)(lds + 2 * base), 16, 0, 0);
                q += kSB;
                row += dq;
                zz += dr;
                if (zz >= BZ) {
                    zz -= BZ;
                    ++row;
                }
            }
        } else {
            // flat element loop (z fastest: coalesced rows), e = row BZ + zz
            // carried from one iteration to the next: e grows by kSB = dq BZ + dr
            // with dr < BZ, so zz + dr < 2 BZ needs at most one carry
            const int dq = kSB / BZ, dr = kSB - dq * BZ;
            int row = (int)threadIdx.x / BZ;
            int zz = (int)threadIdx.x - row * BZ;
            // unrolled: the iterations are independent, so several row-base
            // reads and global loads are in flight before the first wait
            // (measured at 256^3 order 6: fused pass 2.05 -> 1.92 ms)
#ifndef SLC_STAGE_UNROLL
#define SLC_STAGE_UNROLL 4
#endif
#pragma unroll SLC_STAGE_UNROLL
            for (int e = threadIdx.x; e < nbox; e += kSB) {
                lds[e] = S[rowbase[row] + zz];
                row += dq;
                zz += dr;
                if (zz >= BZ) {
                    zz -= BZ;
                    ++row;
                }
            }
        }
    }
    __syncthreads();
    return fits;
}

// z weights of one foot's order-P stencil against the node tables
// (semilag.py:441-454): the part of the weights that reads memory.
template <int P>
struct ZWeights {
    double wz[P];
};

template <int C, int P>
__device__ inline void zweights(const Grid& g, const FootSt& f, ZWeights<P>& w) {
    const double* zn = znodes<C>(g);
    const double* dinv = Comp<C>::centers ? g.dinv_cP : g.dinv_fP;
    z_weights<P>(zn, dinv, f.z, f.bk, w.wz);
}

// Order-P gather of one foot given its z weights: the x and y weights, then
// the staged z -> y -> x contraction with FMAs, from the LDS box when it was
// staged, else from S.
template <typename T, int C, int P>
__device__ inline double contract_w(const Grid& g, const FootSt& f, const ZWeights<P>& w,
                                    const SBox& b, bool fits, const T* __restrict__ lds,
                                    int BY, int BZ, const T* __restrict__ S) {
    double wx[P], wy[P];
    uniform_weights<P>(f.tx, wx);
    uniform_weights<P>(f.ty, wy);
    const double* wz = w.wz;
    double acc = 0.0;
    if (fits) {
        const int ox = f.bx - b.xmin, oy = f.by - b.ymin, oz = f.bk - b.kmin;
#pragma unroll
        for (int i = 0; i < P; ++i) {
            double ai = 0.0;
#pragma unroll
            for (int j = 0; j < P; ++j) {
                const T* col = lds + ((ox + i) * BY + (oy + j)) * BZ + oz;
                double sj = 0.0;
#pragma unroll
                for (int k = 0; k < P; ++k) sj = fma(wz[k], (double)col[k], sj);
                ai = fma(wy[j], sj, ai);
            }
            acc = fma(wx[i], ai, acc);
        }
    } else {
#pragma unroll 1
        for (int i = 0; i < P; ++i) {
            const int gx = wrap_idx(f.bx + i, g.nx);
            double ai = 0.0;
#pragma unroll
            for (int j = 0; j < P; ++j) {
                const int gy = wrap_idx(f.by + j, g.ny);
                const T* col = S + node_addr<C>(g, gx, gy, f.bk);
                double sj = 0.0;
#pragma unroll
                for (int k = 0; k < P; ++k) sj = fma(wz[k], (double)col[k], sj);
                ai = fma(wy[j], sj, ai);
            }
            acc = fma(uniform_weight_m<P>(f.tx, i), ai, acc);
        }
    }
    return acc;
}

// Weights, then the contraction.
template <typename T, int C, int P>
__device__ inline double contract(const Grid& g, const FootSt& f, const SBox& b, bool fits,
                                  const T* __restrict__ lds, int BY, int BZ,
                                  const T* __restrict__ S) {
    ZWeights<P> w;
    zweights<C, P>(g, f, w);
    return contract_w<T, C, P>(g, f, w, b, fits, lds, BY, BZ, S);
}

__device__ inline void box_add(SBox& l, const FootSt& f) {
    l.xmin = min(l.xmin, f.bx); l.xmax = max(l.xmax, f.bx);
    l.ymin = min(l.ymin, f.by); l.ymax = max(l.ymax, f.by);
    l.kmin = min(l.kmin, f.bk); l.kmax = max(l.kmax, f.bk);
}

// Iterated midpoint (semilag.py:534-552) from the arrival point with its
// shared first sample (u0, v0, w0); returns the unclamped foot.
template <int C>
__device__ inline void trace_foot(const Grid& g, const TrajA& V, double xa, double ya,
                                  double za, int K, double u0, double v0, double w0, double h,
                                  double& x, double& y, double& z) {
    double xm = xa - h * u0, ym = ya - h * v0;
    double zm = clampd(za - h * w0, g.z_lo, g.z_hi);
    for (int it = 1; it < g.n_iters; ++it) {
        double us, vs, ws;
        sample3<C>(g, V, xm, ym, zm, za, K, us, vs, ws);
        xm = xa - h * us;
        ym = ya - h * vs;
        zm = clampd(za - h * ws, g.z_lo, g.z_hi);
    }
    if (g.n_iters == 0) { xm = xa; ym = ya; zm = za; }
    x = 2.0 * xm - xa;
    y = 2.0 * ym - ya;
    z = 2.0 * zm - za;
}

// T: storage type of the gathered field (double, or float for the FP32-storage
// variant: values are staged and read as float, weights and accumulation stay
// FP64).  F / Fm are always the FP64 fields (explicit xy-diffusion);
// Sn / Sm are the gathered copies (== F / Fm when T is double).
// One tile of the step kernel (tile = linear tile index, k fastest).  The
// caller provides the block's LDS: lds (LDS_BYTES), sh (6 ints; 12 for
// MODE_BDF2), s_cnt, and for MODE_BDF2 nst (one FootSt per tile point).
// defer (MODE_GATHER): when non-null, one flag per tile: 1 for a tile whose
// box does not fit the LDS budget (it is not gathered here but left for a
// pass with a larger budget), 0 for the others.
// redo (MODE_GATHER): when redo.flags is non-null, only the points whose tile
// of an earlier pass (redo.ti x redo.tj x redo.tk tiles, same numbering) has
// its defer flag set are gathered; the others were done by that pass.
struct Redo {
    const unsigned* flags = nullptr;
    int ti = 1, tj = 1, tk = 1;
};

__device__ inline bool redo_point(const Redo& r, const Grid& g, int nzi, int i, int j, int k) {
    if (!r.flags) return true;
    const int rtk = (nzi + r.tk - 1) / r.tk, rtj = (g.ny + r.tj - 1) / r.tj;
    return r.flags[((long)(i / r.ti) * rtj + j / r.tj) * rtk + k / r.tk] != 0u;
}

template <typename T, int C, int P, int MODE, int TI, int TJ, int TK, int LDS_BYTES>
__device__ __forceinline__ void sl_tile_body(
    const Grid& g, const ArrivalTabs& A, const TrajVel& V, const double* __restrict__ F,
    const double* __restrict__ Fm, const T* __restrict__ Sn, const T* __restrict__ Sm,
    double dt, double* __restrict__ out, const Feet& ft, unsigned long long* nclamp, long tile,
    T* __restrict__ lds, int* sh, int* s_cnt, FootSt* nst = nullptr,
    unsigned* defer = nullptr, Redo redo = Redo{}) {
    constexpr int NPT = TI * TJ * TK;
    static_assert(NPT % kSB == 0, "tile must be a multiple of the block");
    constexpr int PT = NPT / kSB;
    const int nzi = Comp<C>::NZI(g);
    const int ntk = (nzi + TK - 1) / TK, ntj = (g.ny + TJ - 1) / TJ;
    // the step kernels' trajectory velocity is always a stored field
    const TrajA VA{{V.a[0], V.a[1], V.a[2]}};
    const int tk = (int)(tile % ntk);
    const long r0 = tile / ntk;
    const int tj = (int)(r0 % ntj), ti = (int)(r0 / ntj);
    const int i0 = ti * TI, j0 = tj * TJ, k0t = tk * TK;

    // ---- phase 1: feet and stencils --------------------------------------
    long idx[PT];
    FootSt st[PT];
    SBox l{INT_MAX, INT_MIN, INT_MAX, INT_MIN, INT_MAX, INT_MIN};
    SBox ln{INT_MAX, INT_MIN, INT_MAX, INT_MIN, INT_MAX, INT_MIN};  // MODE_BDF2 near box
    int ncl = 0;
#pragma unroll
    for (int p = 0; p < PT; ++p) {
        const int q = threadIdx.x + kSB * p;
        const int kk = q % TK, t2 = q / TK;
        const int jj = t2 % TJ, ii = t2 / TJ;
        const int i = i0 + ii, j = j0 + jj, k = k0t + kk;
        idx[p] = -1;
        if (i < g.nx && j < g.ny && k < nzi &&
            (MODE != MODE_GATHER || redo_point(redo, g, nzi, i, j, k))) {
            const long n = ((long)i * g.ny + j) * nzi + k;
            idx[p] = n;
            double xa, ya, za;
            int I, J, K;
            arrival<C>(g, n, xa, ya, za, I, J, K);
            double x, y, z;
            if (MODE == MODE_NEAR || MODE == MODE_GATHER) {
                x = ft.x[n];
                y = ft.y[n];
                z = ft.z[n];
            } else {
                // the first sample is at the arrival point for every foot depth
                double u0, v0, w0;
                sample3_at(g, VA, A.xdx[I - 1], A.ydy[J - 1], za, A.csel[K - 1], u0, v0, w0);
                if (MODE == MODE_FAR || MODE == MODE_BDF2) {
                    // the near foot first: it leaves the registers (to LDS or
                    // to the feet arrays) before the far foot is traced
                    double xn, yn, zn;
                    trace_foot<C>(g, VA, xa, ya, za, K, u0, v0, w0, 0.5 * dt, xn, yn, zn);
                    if (MODE == MODE_FAR) {
                        ft.x[n] = xn;
                        ft.y[n] = yn;
                        ft.z[n] = zn;
                    } else {
                        FootSt sn;
                        make_st<C, P>(g, xn, yn, zn, sn, ncl);
                        box_add(ln, sn);
                        nst[q] = sn;
                    }
                }
                trace_foot<C>(g, VA, xa, ya, za, K, u0, v0, w0,
                           (MODE == MODE_FAR || MODE == MODE_BDF2) ? 0.5 * (2.0 * dt) : 0.5 * dt,
                           x, y, z);
            }
            make_st<C, P>(g, x, y, z, st[p], ncl);
            box_add(l, st[p]);
        }
    }

    // ---- stage the tile's stencil box in LDS -------------------------------
    // With one point per lane, the z weights are formed before the box is
    // staged, so their table loads are in flight with the staging's.
    constexpr bool HOIST = PT == 1;
    ZWeights<P> w0;
    if (HOIST && idx[0] >= 0) zweights<C, P>(g, st[0], w0);
    const T* __restrict__ S = (MODE == MODE_FAR || MODE == MODE_BDF2) ? Sm : Sn;
    // MODE_BDF2: the near box is reduced with the far one and waits in
    // sh[6..11], so no lane holds it through the far contraction
    SBox b = (MODE == MODE_BDF2) ? s_block_box2(l, ln, sh) : s_block_box(l, sh);
    int BY = 0, BZ = 0;
    const bool fits = stage_box<T, C, P>(g, S, b, lds, LDS_BYTES, BY, BZ);
    if (MODE == MODE_GATHER && defer) {  // fits is uniform over the block
        if (threadIdx.x == 0) defer[tile] = fits ? 0u : 1u;
        if (!fits) return;
    }

    // ---- phase 2: contract -------------------------------------------------
    const double rdx2 = 1.0 / (g.dx * g.dx), rdy2 = 1.0 / (g.dy * g.dy);
    double farv[PT];
    if (MODE == MODE_BDF2) {
        // far value of every point, then the near box into the same buffer
#pragma unroll
        for (int p = 0; p < PT; ++p)
            farv[p] = idx[p] < 0 ? 0.0
                      : HOIST ? contract_w<T, C, P>(g, st[p], w0, b, fits, lds, BY, BZ, S)
                              : contract<T, C, P>(g, st[p], b, fits, lds, BY, BZ, S);
        __syncthreads();  // every lane is done with the far box
        // the near stencils' weights, before the near box is staged
        if (HOIST && idx[0] >= 0) zweights<C, P>(g, nst[threadIdx.x], w0);
    }
    SBox bn = (MODE == MODE_BDF2) ? SBox{sh[6], sh[7], sh[8], sh[9], sh[10], sh[11]} : b;
    const bool fits_n = (MODE == MODE_BDF2) ? stage_box<T, C, P>(g, Sn, bn, lds, LDS_BYTES, BY, BZ)
                                            : fits;
#pragma unroll
    for (int p = 0; p < PT; ++p) {
        if (idx[p] < 0) continue;
        const long n = idx[p];
        const double acc =
            HOIST ? ((MODE == MODE_BDF2)
                         ? contract_w<T, C, P>(g, nst[threadIdx.x], w0, bn, fits_n, lds, BY, BZ,
                                               Sn)
                         : contract_w<T, C, P>(g, st[p], w0, b, fits, lds, BY, BZ, S))
            : (MODE == MODE_BDF2)
                ? contract<T, C, P>(g, nst[threadIdx.x + kSB * p], bn, fits_n, lds, BY, BZ, Sn)
                : contract<T, C, P>(g, st[p], b, fits, lds, BY, BZ, S);
        // arrival ghost index of n
        const unsigned un = (unsigned)n;
        const int K = (int)(un % (unsigned)nzi) + 1;
        const unsigned rr = un / (unsigned)nzi;
        const int J = (int)(rr % (unsigned)g.ny) + 1, I = (int)(rr / (unsigned)g.ny) + 1;
        const long o = ((long)I * Comp<C>::SY(g) + J) * Comp<C>::SZ(g) + K;
        if (MODE == MODE_GATHER) {
            out[n] = acc;
        } else if (MODE == MODE_ADVECT || MODE == MODE_FAR) {
            out[o] = acc;
        } else if (MODE == MODE_BDF1) {
            out[o] = acc + dt * rhs_xy<C>(g, F, I, J, K, rdx2, rdy2);
        } else {
            const double far = (MODE == MODE_BDF2) ? farv[p] : out[o];
            const double rn = rhs_xy<C>(g, F, I, J, K, rdx2, rdy2);
            // Rh: 2R^n - R^{n-1} ('extrap') or R^n ('lagged'), solver.py:848-854
            const double rh = g.xy_lagged ? rn : 2.0 * rn - rhs_xy<C>(g, Fm, I, J, K, rdx2, rdy2);
            const double dt_eff = (2.0 / 3.0) * dt;
            const double pred = fma(far, -1.0 / 3.0, acc * (4.0 / 3.0));
            out[o] = pred + dt_eff * rh;
        }
    }
    if (nclamp) {
        if (threadIdx.x == 0) *s_cnt = 0;
        __syncthreads();
        if (ncl) atomicAdd(s_cnt, ncl);
        __syncthreads();
        if (threadIdx.x == 0 && *s_cnt) atomicAdd(nclamp, (unsigned long long)*s_cnt);
    }
}

}  // namespace slc
