// slc_sl.hip -- semi-Lagrangian kernels for gfx950.
//
//   departure_kernel   feet of every arrival point of one component
//                      (semilag.py:534-552; TritonSL.departure,
//                      semilag_triton.py:384-420)
//   gather_kernel      clamp+count the foot, order-P stencil, gather
//                      (semilag.py:471-510; TritonSL.gather, 422-455)
//   advect_kernel      both fused, writing the ghost-shaped output interior
//                      (SLAdvector.advect, semilag.py:584-671)
//   bdf1_kernel        advect at dt with U* = V^n + dt * nu * L_xy V^n
//                      (solver.py:737-765)
//   bdf2_far/near      two independent feet (dt, 2dt) of the frozen
//                      U* = 2V^n - V^{n-1}, two order-P gathers (V^n at the
//                      near foot, V^{n-1} at the far foot), extrapolated xy
//                      diffusion and the BDF2 predictor (solver.py:846-903)
//
// One thread per arrival point, z fastest so a wavefront's stencils overlap
// in x/y and march along z; blocks are remapped so neighbouring tiles share
// an XCD's L2.  Clamp counts are reduced per block and added with one
// 64-bit atomic per block.
#include "slc_device.h"
#include "slc_internal.h"

namespace slc {

constexpr int kBlock = 256;

__device__ inline void block_count(unsigned long long* counter, int mine) {
    // wavefront ballot-free reduction through LDS; one atomic per block
    __shared__ int s_cnt;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    if (mine) atomicAdd(&s_cnt, mine);
    __syncthreads();
    if (threadIdx.x == 0 && s_cnt) atomicAdd(counter, (unsigned long long)s_cnt);
}

// TV: TrajVel (ghost-shaped trajectory fields) or TrajNodes (packed node
// buffers).  With nclamp the stored z is clamped into [z_lo, z_hi] and the
// clamped feet are counted (TritonSL.departure, semilag_triton.py:262-269);
// without it z is stored unclamped (departure_coords, semilag.py:534-552).
template <int C, int T, typename TV>
__global__ __launch_bounds__(kBlock) void departure_kernel(Grid g, TV V, double dt,
                                                           double* __restrict__ xd,
                                                           double* __restrict__ yd,
                                                           double* __restrict__ zd, long N,
                                                           unsigned long long* nclamp) {
    long nb = (N + kBlock - 1) / kBlock;
    long n = xcd_block(nb) * kBlock + threadIdx.x;
    int clamped = 0;
    if (n < N) {
        double xa, ya, za;
        int I, J, K;
        arrival<C>(g, n, xa, ya, za, I, J, K);
        double x, y, z;
        foot<T, TV>(g, V, xa, ya, za, 0.5 * dt, x, y, z);
        if (nclamp) {
            clamped = (z < g.z_lo || z > g.z_hi) ? 1 : 0;
            z = clampd(z, g.z_lo, g.z_hi);
        }
        xd[n] = x;
        yd[n] = y;
        zd[n] = z;
    }
    if (nclamp) block_count(nclamp, clamped);
}

template <int C, int P>
__global__ __launch_bounds__(kBlock) void gather_kernel(Grid g, const double* __restrict__ F,
                                                        long fsx, long fsy, int fox, int foy,
                                                        const double* __restrict__ xd,
                                                        const double* __restrict__ yd,
                                                        const double* __restrict__ zd,
                                                        double* __restrict__ out, long N,
                                                        unsigned long long* nclamp) {
    long nb = (N + kBlock - 1) / kBlock;
    long n = xcd_block(nb) * kBlock + threadIdx.x;
    int clamped = 0;
    if (n < N) {
        double z = zd[n];
        clamped = (z < g.z_lo || z > g.z_hi) ? 1 : 0;
        z = clampd(z, g.z_lo, g.z_hi);
        Stencil<C, P> s;
        build_stencil<C, P>(g, xd[n], yd[n], z, s);
        if (fsx == 0) {
            out[n] = gather<C, P>(g, F, s);
        } else {
            // caller-described layout (e.g. a packed (NX, NY, NZ) node buffer)
            double acc = 0.0;
#pragma unroll
            for (int i = 0; i < P; ++i) {
                double ai = 0.0;
#pragma unroll
                for (int j = 0; j < P; ++j) {
                    const double* col =
                        F + (long)(s.ix[i] + fox) * fsx + (long)(s.iy[j] + foy) * fsy + s.k0;
                    double sj = 0.0;
#pragma unroll
                    for (int k = 0; k < P; ++k) sj = fma(s.wz[k], col[k], sj);
                    ai = fma(s.wy[j], sj, ai);
                }
                acc = fma(s.wx[i], ai, acc);
            }
            out[n] = acc;
        }
    }
    if (nclamp) block_count(nclamp, clamped);
}

// Stencil base indices of given feet (the index half of _build_iw_impl,
// semilag.py:471-500): x/y node index of stencil offset 0 after the periodic
// wrap, and the clamped z base k0.
template <int C, int P>
__global__ __launch_bounds__(kBlock) void stencil_kernel(Grid g, const double* __restrict__ xd,
                                                         const double* __restrict__ yd,
                                                         const double* __restrict__ zd,
                                                         int* __restrict__ ix,
                                                         int* __restrict__ iy,
                                                         int* __restrict__ k0, long N,
                                                         unsigned long long* nclamp) {
    long n = (long)blockIdx.x * kBlock + threadIdx.x;
    int clamped = 0;
    if (n < N) {
        double z = zd[n];
        clamped = (z < g.z_lo || z > g.z_hi) ? 1 : 0;
        z = clampd(z, g.z_lo, g.z_hi);
        Stencil<C, P> s;
        build_stencil<C, P>(g, xd[n], yd[n], z, s);
        ix[n] = s.ix[0];
        iy[n] = s.iy[0];
        k0[n] = s.k0;
    }
    if (nclamp) block_count(nclamp, clamped);
}

// z-locate diagnostic: node-interval index m0 (semilag.py:413-420) by the
// analytic path and by the fast path used in the kernels.
template <int C>
__global__ void locate_kernel(Grid g, const double* __restrict__ z, long n, int* exact,
                              int* fast) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double zc = clampd(z[i], g.z_lo, g.z_hi);
    exact[i] = locate_from_cell<C>(g, face_cell(g, zc), zc);
    double split;
    const int c = face_cell_fast(g, zc, split);
    fast[i] = locate_split<C>(c, split, zc);
}

hipError_t run_locate(const Grid& g, int comp, const double* z, long n, int* exact, int* fast,
                      hipStream_t s) {
    if (n <= 0) return hipSuccess;
    dim3 grid((unsigned)((n + 255) / 256));
    if (comp == 2) hipLaunchKernelGGL(locate_kernel<2>, grid, dim3(256), 0, s, g, z, n, exact, fast);
    else hipLaunchKernelGGL(locate_kernel<0>, grid, dim3(256), 0, s, g, z, n, exact, fast);
    return hipGetLastError();
}

template <int C>
__device__ inline long out_addr(const Grid& g, int I, int J, int K) {
    return ((long)I * Comp<C>::SY(g) + J) * Comp<C>::SZ(g) + K;
}

template <int C, int P, int T>
__global__ __launch_bounds__(kBlock) void advect_kernel(Grid g, TrajVel V,
                                                        const double* __restrict__ F, double dt,
                                                        double* __restrict__ out, long N,
                                                        unsigned long long* nclamp) {
    long nb = (N + kBlock - 1) / kBlock;
    long n = xcd_block(nb) * kBlock + threadIdx.x;
    int clamped = 0;
    if (n < N) {
        double xa, ya, za;
        int I, J, K;
        arrival<C>(g, n, xa, ya, za, I, J, K);
        double x, y, z;
        foot<T>(g, V, xa, ya, za, 0.5 * dt, x, y, z);
        clamped = (z < g.z_lo || z > g.z_hi) ? 1 : 0;
        z = clampd(z, g.z_lo, g.z_hi);
        Stencil<C, P> s;
        build_stencil<C, P>(g, x, y, z, s);
        out[out_addr<C>(g, I, J, K)] = gather<C, P>(g, F, s);
    }
    if (nclamp) block_count(nclamp, clamped);
}

// BDF1 bootstrap predictor: ustar = advect(V^n; U* = V^n, dt) + dt * R(V^n)
template <int C, int P, int T>
__global__ __launch_bounds__(kBlock) void bdf1_kernel(Grid g, TrajVel V,
                                                      const double* __restrict__ F, double dt,
                                                      double dx2, double dy2,
                                                      double* __restrict__ out, long N,
                                                      unsigned long long* nclamp) {
    long nb = (N + kBlock - 1) / kBlock;
    long n = xcd_block(nb) * kBlock + threadIdx.x;
    int clamped = 0;
    if (n < N) {
        double xa, ya, za;
        int I, J, K;
        arrival<C>(g, n, xa, ya, za, I, J, K);
        double x, y, z;
        foot<T>(g, V, xa, ya, za, 0.5 * dt, x, y, z);
        clamped = (z < g.z_lo || z > g.z_hi) ? 1 : 0;
        z = clampd(z, g.z_lo, g.z_hi);
        Stencil<C, P> s;
        build_stencil<C, P>(g, x, y, z, s);
        double adv = gather<C, P>(g, F, s);
        double r = rhs_xy<C>(g, F, I, J, K, dx2, dy2);
        out[out_addr<C>(g, I, J, K)] = adv + dt * r;
    }
    if (nclamp) block_count(nclamp, clamped);
}

// BDF2 predictor (noninc pressure, extrapolated xy RHS):
//   u_hat = (4/3) ubar(dt; V^n) - (1/3) ubarbar(2dt; V^{n-1}) + (2dt/3)(2R^n - R^{n-1})
// Two passes over the arrival points keep each pass at one order-P gather:
//   far:  out <- V^{n-1} at the foot of depth 2dt          (solver.py:872-882)
//   near: out <- 4/3 V^n(foot dt) - 1/3 out + dt_eff Rh    (solver.py:885-903)
template <int C, int P, int T>
__global__ __launch_bounds__(kBlock) void bdf2_far_kernel(Grid g, TrajVel V,
                                                          const double* __restrict__ Fm,
                                                          double dt, double* __restrict__ out,
                                                          long N, unsigned long long* nclamp) {
    long nb = (N + kBlock - 1) / kBlock;
    long n = xcd_block(nb) * kBlock + threadIdx.x;
    int clamped = 0;
    if (n < N) {
        double xa, ya, za;
        int I, J, K;
        arrival<C>(g, n, xa, ya, za, I, J, K);
        double x, y, z;
        foot<T>(g, V, xa, ya, za, 0.5 * (2.0 * dt), x, y, z);
        clamped = (z < g.z_lo || z > g.z_hi) ? 1 : 0;
        z = clampd(z, g.z_lo, g.z_hi);
        Stencil<C, P> s;
        build_stencil<C, P>(g, x, y, z, s);
        out[out_addr<C>(g, I, J, K)] = gather<C, P>(g, Fm, s);
    }
    if (nclamp) block_count(nclamp, clamped);
}

template <int C, int P, int T>
__global__ __launch_bounds__(kBlock) void bdf2_near_kernel(Grid g, TrajVel V,
                                                           const double* __restrict__ Fn,
                                                           const double* __restrict__ Fm,
                                                           double dt, double dt_eff, double dx2,
                                                           double dy2, double* __restrict__ out,
                                                           long N, unsigned long long* nclamp) {
    long nb = (N + kBlock - 1) / kBlock;
    long n = xcd_block(nb) * kBlock + threadIdx.x;
    int clamped = 0;
    if (n < N) {
        double xa, ya, za;
        int I, J, K;
        arrival<C>(g, n, xa, ya, za, I, J, K);
        double x, y, z;
        foot<T>(g, V, xa, ya, za, 0.5 * dt, x, y, z);
        clamped = (z < g.z_lo || z > g.z_hi) ? 1 : 0;
        z = clampd(z, g.z_lo, g.z_hi);
        Stencil<C, P> s;
        build_stencil<C, P>(g, x, y, z, s);
        double near = gather<C, P>(g, Fn, s);
        const long o = out_addr<C>(g, I, J, K);
        double far = out[o];
        double rn = rhs_xy<C>(g, Fn, I, J, K, dx2, dy2);
        // Rh: 2R^n - R^{n-1} ('extrap') or R^n ('lagged'), solver.py:848-854
        double rh = g.xy_lagged ? rn : 2.0 * rn - rhs_xy<C>(g, Fm, I, J, K, dx2, dy2);
        double pred = fma(far, -1.0 / 3.0, near * (4.0 / 3.0));
        out[o] = pred + dt_eff * rh;
    }
    if (nclamp) block_count(nclamp, clamped);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

static inline long n_arrival(const Grid& g, int comp) {
    return (long)g.nx * g.ny * (comp == 2 ? g.nz - 1 : g.nz);
}

static inline dim3 grid_for(long N) { return dim3((unsigned)((N + kBlock - 1) / kBlock)); }

// Launch KER<C, P, T> for the plan's field order P and trajectory order T.
#define SLC_LAUNCH_PT(KER, ...)                                                    \
    do {                                                                           \
        if (g.order == 4 && g.traj_order == 4)                                     \
            hipLaunchKernelGGL((KER<C, 4, 4>), __VA_ARGS__);                       \
        else if (g.order == 4)                                                     \
            hipLaunchKernelGGL((KER<C, 4, 2>), __VA_ARGS__);                       \
        else if (g.traj_order == 4)                                                \
            hipLaunchKernelGGL((KER<C, 6, 4>), __VA_ARGS__);                       \
        else                                                                       \
            hipLaunchKernelGGL((KER<C, 6, 2>), __VA_ARGS__);                       \
    } while (0)

#define SLC_DISPATCH_C(comp, FN, ...) \
    switch (comp) {                     \
        case 0: FN<0>(__VA_ARGS__); break; \
        case 1: FN<1>(__VA_ARGS__); break; \
        default: FN<2>(__VA_ARGS__); break; \
    }

template <int C, typename TV>
static void launch_departure_tv(const Grid& g, const TV& V, double dt, double* xd, double* yd,
                                double* zd, unsigned long long* ncl, hipStream_t s) {
    long N = n_arrival(g, C);
    if (N <= 0) return;
    if (g.traj_order == 4)
        hipLaunchKernelGGL((departure_kernel<C, 4, TV>), grid_for(N), dim3(kBlock), 0, s, g, V,
                           dt, xd, yd, zd, N, ncl);
    else
        hipLaunchKernelGGL((departure_kernel<C, 2, TV>), grid_for(N), dim3(kBlock), 0, s, g, V,
                           dt, xd, yd, zd, N, ncl);
}

template <int C>
static void launch_departure(const Grid& g, const TrajVel& V, int layout, double dt, double* xd,
                             double* yd, double* zd, unsigned long long* ncl, hipStream_t s) {
    if (layout == 1)
        launch_departure_tv<C, TrajNodes>(g, TrajNodes{{V.a[0], V.a[1], V.a[2]}}, dt, xd, yd, zd,
                                          ncl, s);
    else
        launch_departure_tv<C, TrajVel>(g, V, dt, xd, yd, zd, ncl, s);
}

hipError_t run_departure(const Grid& g, int comp, const TrajVel& V, int layout, double dt,
                         double* xd, double* yd, double* zd, unsigned long long* ncl,
                         hipStream_t s) {
    SLC_DISPATCH_C(comp, launch_departure, g, V, layout, dt, xd, yd, zd, ncl, s);
    return hipGetLastError();
}

template <int C, int P>
static void launch_gather_cp(const Grid& g, const double* F, long fsx, long fsy, int fox, int foy,
                             const double* xd, const double* yd, const double* zd, double* out,
                             unsigned long long* ncl, hipStream_t s) {
    long N = n_arrival(g, C);
    if (N <= 0) return;
    hipLaunchKernelGGL((gather_kernel<C, P>), grid_for(N), dim3(kBlock), 0, s, g, F, fsx, fsy,
                       fox, foy, xd, yd, zd, out, N, ncl);
}

template <int C>
static void launch_gather(const Grid& g, const double* F, long fsx, long fsy, int fox, int foy,
                          const double* xd, const double* yd, const double* zd, double* out,
                          unsigned long long* ncl, hipStream_t s) {
    if (g.order == 4)
        launch_gather_cp<C, 4>(g, F, fsx, fsy, fox, foy, xd, yd, zd, out, ncl, s);
    else
        launch_gather_cp<C, 6>(g, F, fsx, fsy, fox, foy, xd, yd, zd, out, ncl, s);
}

hipError_t run_gather(const Grid& g, int comp, const double* F, long fsx, long fsy, int fox,
                      int foy, const double* xd, const double* yd, const double* zd, double* out,
                      unsigned long long* ncl, hipStream_t s) {
    SLC_DISPATCH_C(comp, launch_gather, g, F, fsx, fsy, fox, foy, xd, yd, zd, out, ncl, s);
    return hipGetLastError();
}

template <int C>
static void launch_stencil(const Grid& g, const double* xd, const double* yd, const double* zd,
                           int* ix, int* iy, int* k0, unsigned long long* ncl, hipStream_t s) {
    long N = n_arrival(g, C);
    if (N <= 0) return;
    if (g.order == 4)
        hipLaunchKernelGGL((stencil_kernel<C, 4>), grid_for(N), dim3(kBlock), 0, s, g, xd, yd, zd,
                           ix, iy, k0, N, ncl);
    else
        hipLaunchKernelGGL((stencil_kernel<C, 6>), grid_for(N), dim3(kBlock), 0, s, g, xd, yd, zd,
                           ix, iy, k0, N, ncl);
}

hipError_t run_stencil(const Grid& g, int comp, const double* xd, const double* yd,
                       const double* zd, int* ix, int* iy, int* k0, unsigned long long* ncl,
                       hipStream_t s) {
    SLC_DISPATCH_C(comp, launch_stencil, g, xd, yd, zd, ix, iy, k0, ncl, s);
    return hipGetLastError();
}

template <int C>
static void launch_advect(const Grid& g, const TrajVel& V, const double* F, double dt,
                          double* out, unsigned long long* ncl, hipStream_t s) {
    long N = n_arrival(g, C);
    if (N <= 0) return;
    SLC_LAUNCH_PT(advect_kernel, grid_for(N), dim3(kBlock), 0, s, g, V, F, dt, out, N, ncl);
}

hipError_t run_advect(const Grid& g, int comp, const TrajVel& V, const double* F, double dt,
                      double* out, unsigned long long* ncl, hipStream_t s) {
    SLC_DISPATCH_C(comp, launch_advect, g, V, F, dt, out, ncl, s);
    return hipGetLastError();
}

template <int C>
static void launch_bdf1(const Grid& g, const TrajVel& V, const double* F, double dt, double dx2,
                        double dy2, double* out, unsigned long long* ncl, hipStream_t s) {
    long N = n_arrival(g, C);
    if (N <= 0) return;
    SLC_LAUNCH_PT(bdf1_kernel, grid_for(N), dim3(kBlock), 0, s, g, V, F, dt, dx2, dy2, out, N,
                  ncl);
}

hipError_t run_bdf1(const Grid& g, int comp, const TrajVel& V, const double* F, double dt,
                    double* out, unsigned long long* ncl, hipStream_t s) {
    // reciprocal squares (rhs_xy multiplies)
    double dx2 = 1.0 / (g.dx * g.dx), dy2 = 1.0 / (g.dy * g.dy);
    SLC_DISPATCH_C(comp, launch_bdf1, g, V, F, dt, dx2, dy2, out, ncl, s);
    return hipGetLastError();
}

template <int C>
static void launch_bdf2(const Grid& g, int pass, const TrajVel& V, const double* Fn,
                        const double* Fm, double dt, double dt_eff, double dx2, double dy2,
                        double* out, unsigned long long* ncl, hipStream_t s) {
    long N = n_arrival(g, C);
    if (N <= 0) return;
    if (pass == 0) {
        SLC_LAUNCH_PT(bdf2_far_kernel, grid_for(N), dim3(kBlock), 0, s, g, V, Fm, dt, out, N,
                      ncl);
    } else {
        SLC_LAUNCH_PT(bdf2_near_kernel, grid_for(N), dim3(kBlock), 0, s, g, V, Fn, Fm, dt,
                      dt_eff, dx2, dy2, out, N, ncl);
    }
}

// pass 0: far foot (V^{n-1} at depth 2dt into out); pass 1: near foot + predictor
hipError_t run_bdf2(const Grid& g, int comp, int pass, const TrajVel& V, const double* Fn,
                    const double* Fm, double dt, double* out, unsigned long long* ncl,
                    hipStream_t s) {
    // reciprocal squares (rhs_xy multiplies)
    double dx2 = 1.0 / (g.dx * g.dx), dy2 = 1.0 / (g.dy * g.dy);
    double dt_eff = (2.0 / 3.0) * dt;
    SLC_DISPATCH_C(comp, launch_bdf2, g, pass, V, Fn, Fm, dt, dt_eff, dx2, dy2, out, ncl, s);
    return hipGetLastError();
}

}  // namespace slc
