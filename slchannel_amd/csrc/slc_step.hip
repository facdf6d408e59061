// slc_step.hip -- LDS-tiled semi-Lagrangian step kernels for gfx950.
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
// A workgroup owns a TI x TJ x TK tile of arrival points (z fastest).  Phase 1
// traces the tile's feet and keeps each foot's stencil (base indices,
// fractional offsets, clamped z) in registers; the tile's stencil bounding
// box (its trajectory-CFL halo, measured) is reduced and staged in LDS with
// coalesced z-contiguous loads; phase 2 contracts every P^3 stencil from LDS
// (z, then y, then x, with FMAs -- the same order as the direct gather).  A
// box larger than the LDS budget falls back to global loads for that tile.
#include <climits>

#include "slc_device.h"
#include "slc_internal.h"

namespace slc {

constexpr int kSB = 256;  // threads per workgroup

enum { MODE_ADVECT = 0, MODE_BDF1 = 1, MODE_FAR = 2, MODE_NEAR = 3, MODE_GATHER = 4 };

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

__device__ inline int s_wave_min(int v) {
    for (int o = 32; o > 0; o >>= 1) {
        int w = __shfl_xor(v, o);
        v = w < v ? w : v;
    }
    return v;
}

__device__ inline int s_wave_max(int v) {
    for (int o = 32; o > 0; o >>= 1) {
        int w = __shfl_xor(v, o);
        v = w > v ? w : v;
    }
    return v;
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
// identical results).
__device__ inline void sample3(const Grid& g, const TrajVel& V, double x, double y, double z,
                               double& us, double& vs, double& ws) {
    const double zc = clampd(z, g.z_lo, g.z_hi);
    const double xdx = x / g.dx, ydy = y / g.dy;
    const int cell = face_cell_fast(g, zc);
    us = trilinear<0>(g, V, xdx, ydy, zc, cell);
    vs = trilinear<1>(g, V, xdx, ydy, zc, cell);
    ws = trilinear<2>(g, V, xdx, ydy, zc, cell);
}

// Arrival-point trajectory sample with precomputed x/dx, y/dy and face cell
// (the same values sample3 computes; see ArrivalTabs).
__device__ inline void sample3_at(const Grid& g, const TrajVel& V, double xdx, double ydy,
                                  double z, int cell, double& us, double& vs, double& ws) {
    const double zc = clampd(z, g.z_lo, g.z_hi);
    us = trilinear<0>(g, V, xdx, ydy, zc, cell);
    vs = trilinear<1>(g, V, xdx, ydy, zc, cell);
    ws = trilinear<2>(g, V, xdx, ydy, zc, cell);
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

// T: storage type of the gathered field (double, or float for the FP32-storage
// variant: values are staged and read as float, weights and accumulation stay
// FP64).  F / Fm are always the FP64 fields (explicit xy-diffusion);
// Sn / Sm are the gathered copies (== F / Fm when T is double).
template <typename T, int C, int P, int MODE, int TI, int TJ, int TK, int LDS_BYTES>
__global__ __launch_bounds__(kSB) void sl_tile_kernel(
    Grid g, ArrivalTabs A, TrajVel V, const double* __restrict__ F,
    const double* __restrict__ Fm, const T* __restrict__ Sn, const T* __restrict__ Sm,
    double dt, double* __restrict__ out, Feet ft, unsigned long long* nclamp) {
    constexpr int NPT = TI * TJ * TK;
    static_assert(NPT % kSB == 0, "tile must be a multiple of the block");
    constexpr int PT = NPT / kSB;
    constexpr int CAP = LDS_BYTES / (int)sizeof(T);
    __shared__ T lds[CAP];
    __shared__ int sh[6];
    __shared__ int s_cnt;
    const int nzi = Comp<C>::NZI(g);
    const int ntk = (nzi + TK - 1) / TK, ntj = (g.ny + TJ - 1) / TJ;
    const long tile = xcd_block(gridDim.x);
    const int tk = (int)(tile % ntk);
    const long r0 = tile / ntk;
    const int tj = (int)(r0 % ntj), ti = (int)(r0 / ntj);
    const int i0 = ti * TI, j0 = tj * TJ, k0t = tk * TK;

    // ---- phase 1: feet and stencils --------------------------------------
    long idx[PT];
    FootSt st[PT];
    SBox l{INT_MAX, INT_MIN, INT_MAX, INT_MIN, INT_MAX, INT_MIN};
    int ncl = 0;
#pragma unroll
    for (int p = 0; p < PT; ++p) {
        const int q = threadIdx.x + kSB * p;
        const int kk = q % TK, t2 = q / TK;
        const int jj = t2 % TJ, ii = t2 / TJ;
        const int i = i0 + ii, j = j0 + jj, k = k0t + kk;
        idx[p] = -1;
        if (i < g.nx && j < g.ny && k < nzi) {
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
                // iterated midpoint (semilag.py:534-552); the first sample is
                // at the arrival point for every foot depth
                double u0, v0, w0;
                sample3_at(g, V, A.xdx[I - 1], A.ydy[J - 1], za, A.cell[K - 1], u0, v0, w0);
                const double hf = (MODE == MODE_FAR) ? 0.5 * (2.0 * dt) : 0.5 * dt;
                double xm = xa - hf * u0, ym = ya - hf * v0;
                double zm = clampd(za - hf * w0, g.z_lo, g.z_hi);
                for (int it = 1; it < g.n_iters; ++it) {
                    double us, vs, ws;
                    sample3(g, V, xm, ym, zm, us, vs, ws);
                    xm = xa - hf * us;
                    ym = ya - hf * vs;
                    zm = clampd(za - hf * ws, g.z_lo, g.z_hi);
                }
                if (g.n_iters == 0) { xm = xa; ym = ya; zm = za; }
                x = 2.0 * xm - xa;
                y = 2.0 * ym - ya;
                z = 2.0 * zm - za;
                if (MODE == MODE_FAR) {
                    const double hn = 0.5 * dt;
                    double xn = xa - hn * u0, yn = ya - hn * v0;
                    double zn = clampd(za - hn * w0, g.z_lo, g.z_hi);
                    for (int it = 1; it < g.n_iters; ++it) {
                        double us, vs, ws;
                        sample3(g, V, xn, yn, zn, us, vs, ws);
                        xn = xa - hn * us;
                        yn = ya - hn * vs;
                        zn = clampd(za - hn * ws, g.z_lo, g.z_hi);
                    }
                    if (g.n_iters == 0) { xn = xa; yn = ya; zn = za; }
                    ft.x[n] = 2.0 * xn - xa;
                    ft.y[n] = 2.0 * yn - ya;
                    ft.z[n] = 2.0 * zn - za;
                }
            }
            make_st<C, P>(g, x, y, z, st[p], ncl);
            l.xmin = min(l.xmin, st[p].bx); l.xmax = max(l.xmax, st[p].bx);
            l.ymin = min(l.ymin, st[p].by); l.ymax = max(l.ymax, st[p].by);
            l.kmin = min(l.kmin, st[p].bk); l.kmax = max(l.kmax, st[p].bk);
        }
    }

    // ---- stage the tile's stencil box in LDS -------------------------------
    const T* __restrict__ S = (MODE == MODE_FAR) ? Sm : Sn;
    const SBox b = s_block_box(l, sh);
    const int BX = b.xmax - b.xmin + P, BY = b.ymax - b.ymin + P, BZ = b.kmax - b.kmin + P;
    const bool fits = b.xmin <= b.xmax && (long)BX * BY * BZ <= CAP;
    if (fits) {
        // flat element loop (z fastest: coalesced rows); the quotients e/BZ
        // and row/BY come from float reciprocals -- exact here because
        // e < CAP < 2^13 keeps every fractional part >= 0.5/BZ from an integer
        const int rows = BX * BY;
        const float rbz = 1.0f / (float)BZ, rby = 1.0f / (float)BY;
        for (int e = threadIdx.x; e < rows * BZ; e += kSB) {
            int row = (int)(((float)e + 0.5f) * rbz);
            row += (row + 1) * BZ <= e ? 1 : 0;  // exact even if the reciprocal
            row -= row * BZ > e ? 1 : 0;         // were off by an ulp
            const int zz = e - row * BZ;
            int x = (int)(((float)row + 0.5f) * rby);
            x += (x + 1) * BY <= row ? 1 : 0;
            x -= x * BY > row ? 1 : 0;
            const int y = row - x * BY;
            const int gx = wrap_idx(b.xmin + x, g.nx), gy = wrap_idx(b.ymin + y, g.ny);
            lds[e] = S[node_addr<C>(g, gx, gy, b.kmin + zz)];
        }
    }
    __syncthreads();

    // ---- phase 2: contract -------------------------------------------------
    const double* zn = znodes<C>(g);
    const double* dinv = Comp<C>::centers ? g.dinv_cP : g.dinv_fP;
    const double dx2 = g.dx * g.dx, dy2 = g.dy * g.dy;
#pragma unroll
    for (int p = 0; p < PT; ++p) {
        if (idx[p] < 0) continue;
        const long n = idx[p];
        const FootSt& f = st[p];
        double wx[P], wy[P], wz[P];
        uniform_weights<P>(f.tx, wx);
        uniform_weights<P>(f.ty, wy);
        z_weights<P>(zn, dinv, f.z, f.bk, wz);
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
            out[o] = acc + dt * rhs_xy<C>(g, F, I, J, K, dx2, dy2);
        } else {
            const double far = out[o];
            const double rn = rhs_xy<C>(g, F, I, J, K, dx2, dy2);
            const double rm = rhs_xy<C>(g, Fm, I, J, K, dx2, dy2);
            const double rh = 2.0 * rn - rm;
            const double dt_eff = (2.0 / 3.0) * dt;
            const double pred = fma(far, -1.0 / 3.0, acc * (4.0 / 3.0));
            out[o] = pred + dt_eff * rh;
        }
    }
    if (nclamp) {
        if (threadIdx.x == 0) s_cnt = 0;
        __syncthreads();
        if (ncl) atomicAdd(&s_cnt, ncl);
        __syncthreads();
        if (threadIdx.x == 0 && s_cnt) atomicAdd(nclamp, (unsigned long long)s_cnt);
    }
}

// Arrival tables of component C: x/dx, y/dy of the arrival columns and the
// face cell of each arrival z, evaluated with the very functions the
// trajectory sampling uses (so looked-up values are the computed ones).
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
        cell[t] = face_cell(g, clampd(za, g.z_lo, g.z_hi));
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

// tile shape of the step kernels: 4 x 4 x 32 arrival points, 40 KB of LDS
constexpr int kTI = 4, kTJ = 4, kTK = 32, kLDS = 40960;

struct StepArgs {
    const double* F;
    const double* Fm;
    const void* Sn;  // gathered copies (double or float)
    const void* Sm;
    double dt;
    double* out;
    Feet ft;
    unsigned long long* ncl;
};

template <typename T, int C, int P, int MODE>
static void launch_mode(const Grid& g, const ArrivalTabs& A, const TrajVel& V,
                        const StepArgs& a, hipStream_t s) {
    const long nzi = (C == 2) ? g.nz - 1 : g.nz;
    const unsigned tiles = (unsigned)(((long)(g.nx + kTI - 1) / kTI) *
                                      ((g.ny + kTJ - 1) / kTJ) * ((nzi + kTK - 1) / kTK));
    if (tiles == 0) return;
    hipLaunchKernelGGL((sl_tile_kernel<T, C, P, MODE, kTI, kTJ, kTK, kLDS>), dim3(tiles),
                       dim3(kSB), 0, s, g, A, V, a.F, a.Fm, (const T*)a.Sn, (const T*)a.Sm, a.dt,
                       a.out, a.ft, a.ncl);
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
        default: launch_c<T, MODE_GATHER>(g, comp, A, V, a, s); break;
    }
}

hipError_t run_tile_step(const Grid& g, int comp, int mode, const ArrivalTabs& A,
                         const TrajVel& V, const double* F, const double* Fm, const float* Fn32,
                         const float* Fm32, double dt, double* out, double* fx, double* fy,
                         double* fz, unsigned long long* ncl, hipStream_t s) {
    const Feet ft{fx, fy, fz};
    if (Fn32 || Fm32) {
        StepArgs a{F, Fm, Fn32, Fm32, dt, out, ft, ncl};
        launch_t<float>(g, comp, mode, A, V, a, s);
    } else {
        StepArgs a{F, Fm, F, Fm, dt, out, ft, ncl};
        launch_t<double>(g, comp, mode, A, V, a, s);
    }
    return hipGetLastError();
}

}  // namespace slc
