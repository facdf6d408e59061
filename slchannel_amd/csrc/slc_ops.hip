// slc_ops.hip -- field operators of the SL-BDF2 step for gfx950:
//   ghost/boundary fill            solver.py:70-107, projection.py:180-195
//   theta-method z-diffusion       operators.py:382-630 (Thomas, shared factor)
//   divergence (+transpose)        utils.py:166-180
//   per-wavenumber Poisson Thomas  projection.py:133-247 (coefficients on the fly)
//   gradient projection            projection.py:250-272
//   bulk-flux constraint           solver.py:694-723, utils.py:147-163
// (device bodies in slc_ops_dev.h)
//
// All of these are HBM-bound (well under one flop per byte); they are written
// for coalesced 64-lane access: z-contiguous fields are read/written along z,
// the spectral data lives plane-major (k, i, j) so rocFFT transforms contiguous
// x-y planes and the Thomas sweep over k reads consecutive modes per lane.
// Every pass is deterministic (fixed partition, fixed reduction trees).
#include "slc_ops_dev.h"

namespace slc {

// ---------------------------------------------------------------------------
// fills (graph-capture friendly replacements for memset)
// ---------------------------------------------------------------------------

__global__ void fill_u64_kernel(unsigned long long* p, long n, unsigned long long v) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void fill_f64_kernel(double* p, long n, double v) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t run_fill_u64(unsigned long long* p, long n, unsigned long long v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n,
                       v);
    return hipGetLastError();
}

hipError_t run_fill_f64(double* p, long n, double v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_f64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n,
                       v);
    return hipGetLastError();
}

__global__ void axpby_kernel(double* out, const double* a, const double* b, double alpha,
                             double beta, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = alpha * a[i] + beta * b[i];
}

hipError_t run_axpby(double* out, const double* a, const double* b, double alpha, double beta,
                     long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(axpby_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, out, a, b,
                       alpha, beta, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// ghost fill
// ---------------------------------------------------------------------------

__global__ void bc_kernel(Grid g, double* f0, double* f1, double* f2, int kind0) {
    const int kind = kind0 == 3 ? 3 : (int)blockIdx.y;
    double* F = kind0 == 3 ? f0 : (blockIdx.y == 0 ? f0 : (blockIdx.y == 1 ? f1 : f2));
    if (!F) return;
    bc_body(g, F, kind, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

hipError_t run_bc(const Grid& g, double* u, double* v, double* w, hipStream_t st) {
    long nmax = 0;
    for (int k = 0; k < 3; ++k) {
        long n = ghost_count(k, g.nx, g.ny, g.nz);
        if (n > nmax) nmax = n;
    }
    hipLaunchKernelGGL(bc_kernel, dim3((unsigned)((nmax + 255) / 256), 3), dim3(256), 0, st, g, u,
                       v, w, 0);
    return hipGetLastError();
}

hipError_t run_bc_p(const Grid& g, double* p, hipStream_t st) {
    long n = ghost_count(3, g.nx, g.ny, g.nz);
    hipLaunchKernelGGL(bc_kernel, dim3((unsigned)((n + 255) / 256), 1), dim3(256), 0, st, g, p,
                       nullptr, nullptr, 3);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// implicit z-diffusion
// ---------------------------------------------------------------------------
// The shared factorisation is built by one thread (n is at most a few
// thousand).

__global__ void zfactor_kernel(Grid g, int comp, double dt, double theta, const double* dz_c,
                               const double* dz_f, double* tab) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int nz = g.nz;
    const int n = comp == 2 ? nz - 1 : nz;
    double* A = tab;
    double* CP = tab + n;
    double* INV = tab + 2 * n;
    const double alpha = theta * dt * g.nu;
    double cprev = 0.0;
    for (int k = 0; k < n; ++k) {
        double dzl, dzr, dzc;
        if (comp < 2) { dzl = dz_c[k]; dzr = dz_c[k + 1]; dzc = dz_f[k]; }
        else { dzl = dz_f[k]; dzr = dz_f[k + 1]; dzc = dz_c[k + 1]; }
        double lower = -alpha / (dzl * dzc);
        double center = 1.0 + alpha * (1.0 / dzl + 1.0 / dzr) / dzc;
        double upper = -alpha / (dzr * dzc);
        double a = lower, b = center, c = upper;
        if (k == 0) {
            a = 0.0;
            if (comp < 2) b = center - lower;
        }
        if (k == n - 1) {
            if (comp < 2) b = g.top_neumann ? center + upper : center - upper;
            c = 0.0;
        }
        // (k == 0 and k == n-1 coincide only for n == 1)
        double den = (k == 0) ? b : b - a * cprev;
        double cp = c / den;
        A[k] = a;
        CP[k] = cp;
        INV[k] = 1.0 / den;
        cprev = cp;
    }
}

hipError_t run_zfactor(const Grid& g, int comp, double dt, double theta, const double* dz_c,
                       const double* dz_f, double* tab, hipStream_t s) {
    hipLaunchKernelGGL(zfactor_kernel, dim3(1), dim3(64), 0, s, g, comp, dt, theta, dz_c, dz_f,
                       tab);
    return hipGetLastError();
}

template <int C>
__global__ __launch_bounds__(256) void zsolve_kernel(Grid g, double* __restrict__ F,
                                                     const double* __restrict__ tab, double beta,
                                                     const double* __restrict__ dz_c,
                                                     const double* __restrict__ dz_f,
                                                     int kZCols) {
    extern __shared__ double lds[];
    __shared__ long cbase[kZColsMax];
    zsolve_body<C>(g, F, tab, beta, dz_c, dz_f, kZCols, blockIdx.x, lds, cbase);
}

// columns per wavefront of the single-run z-solve (compile-time sweep knob)
#ifndef SLC_ZSCAN_K
#define SLC_ZSCAN_K 1
#endif
constexpr int kZScanK = SLC_ZSCAN_K;
// Columns of 8 or more unknowns per lane move through LDS (coalesced), the
// shorter ones straight to the lanes.  Measured: 1024x1024x512 (8 per lane)
// 4.1 -> 2.8 ms per component; 256^3 (4 per lane) 0.0825 -> 0.0845 ms.
#ifndef SLC_ZSCAN_LDS_MIN_CH
#define SLC_ZSCAN_LDS_MIN_CH 8
#endif

template <int C, int CH, int K>
__global__ __launch_bounds__(256) void zsolve_scan_kernel(Grid g, double* __restrict__ F,
                                                          const double* __restrict__ tab,
                                                          double beta,
                                                          const double* __restrict__ dz_c,
                                                          const double* __restrict__ dz_f) {
    constexpr bool X = CH >= SLC_ZSCAN_LDS_MIN_CH;
    __shared__ double xl[kZScanCols][X ? zscan_lds<CH>(K) : 1];
    const long ncols = (long)g.nx * g.ny;
    const long col0 = ((long)blockIdx.x * kZScanCols + (threadIdx.x >> 6)) * K;
    if (col0 < ncols)
        zsolve_scan_cols<C, CH, K>(g, F, tab, col0, ncols, beta, dz_c, dz_f,
                                   X ? xl[threadIdx.x >> 6] : nullptr);
}

template <int C>
void launch_zsolve_scan(const Grid& g, double* f, const double* tab, int ch, double beta,
                        const double* dz_c, const double* dz_f, hipStream_t s) {
    const long ncols = (long)g.nx * g.ny;
    // one column per wavefront (a single small grid needs the workgroups: four
    // short columns per wavefront, as the ensemble solves them, measured
    // 14.8 -> 17.3 us at 96x64x128); 16 unknowns per lane: two would not fit
    // the registers
    const long K = ch <= 8 ? kZScanK : 1;
    const long per = (long)kZScanCols * K;
    const dim3 grid((unsigned)((ncols + per - 1) / per)), blk(256);
#define SLC_ZSCAN(CHV)                                                                        \
    hipLaunchKernelGGL((zsolve_scan_kernel<C, CHV, (CHV <= 8 ? kZScanK : 1)>), grid, blk, 0, s, \
                       g, f, tab, beta, dz_c, dz_f)
    switch (ch) {
        case 1: SLC_ZSCAN(1); break;
        case 2: SLC_ZSCAN(2); break;
        case 4: SLC_ZSCAN(4); break;
        case 8: SLC_ZSCAN(8); break;
        default: SLC_ZSCAN(16); break;
    }
#undef SLC_ZSCAN
}

hipError_t run_zsolve(const Grid& g, int comp, double* f, double dt, double theta,
                      const double* dz_c, const double* dz_f, double* tab, bool refactor,
                      hipStream_t s) {
    if (refactor)
        hipLaunchKernelGGL(zfactor_kernel, dim3(1), dim3(64), 0, s, g, comp, dt, theta, dz_c,
                           dz_f, tab);
    double beta = theta < 1.0 ? (1.0 - theta) * dt * g.nu : 0.0;
    const int ch = zsolve_scan_ch(comp, g.nz);
    if (ch > 0) {
        if (comp == 0) launch_zsolve_scan<0>(g, f, tab, ch, beta, dz_c, dz_f, s);
        else if (comp == 1) launch_zsolve_scan<1>(g, f, tab, ch, beta, dz_c, dz_f, s);
        else launch_zsolve_scan<2>(g, f, tab, ch, beta, dz_c, dz_f, s);
        return hipGetLastError();
    }
    long ncols = (long)g.nx * g.ny;
    const int cols = zsolve_cols(comp, g.nz);
    if (cols < 1) return hipErrorInvalidValue;  // nz beyond ~6000: column does not fit LDS
    dim3 grid((unsigned)((ncols + cols - 1) / cols));
    size_t shm = (size_t)cols * zsolve_ld(comp, g.nz) * sizeof(double);
    if (comp == 0)
        hipLaunchKernelGGL(zsolve_kernel<0>, grid, dim3(256), shm, s, g, f, tab, beta, dz_c, dz_f,
                           cols);
    else if (comp == 1)
        hipLaunchKernelGGL(zsolve_kernel<1>, grid, dim3(256), shm, s, g, f, tab, beta, dz_c, dz_f,
                           cols);
    else
        hipLaunchKernelGGL(zsolve_kernel<2>, grid, dim3(256), shm, s, g, f, tab, beta, dz_c, dz_f,
                           cols);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// divergence -> plane-major (k, i, j) Poisson right-hand side
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void div_planes_kernel(Grid g, const double* __restrict__ u,
                                                         const double* __restrict__ v,
                                                         const double* __restrict__ w,
                                                         const double* __restrict__ dz_f,
                                                         double scale,
                                                         double* __restrict__ planes) {
    __shared__ double tile[kT][kT + 1];
    div_planes_body(g, u, v, w, dz_f, scale, planes, blockIdx.z, blockIdx.y * kT,
                    blockIdx.x * kT, tile);
}

hipError_t run_div_planes(const Grid& g, const double* u, const double* v, const double* w,
                          const double* dz_f, double scale, double* planes, hipStream_t s) {
    dim3 grid((g.nz + kT - 1) / kT, (g.ny + kT - 1) / kT, g.nx);
    hipLaunchKernelGGL(div_planes_kernel, grid, dim3(256), 0, s, g, u, v, w, dz_f, scale, planes);
    return hipGetLastError();
}

__global__ void div_field_kernel(Grid g, const double* __restrict__ u,
                                 const double* __restrict__ v, const double* __restrict__ w,
                                 const double* __restrict__ dz_f, double* __restrict__ div) {
    long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long N = (long)g.nx * g.ny * g.nz;
    if (n >= N) return;
    const unsigned un = (unsigned)n;
    int k = (int)(un % (unsigned)g.nz);
    unsigned r = un / (unsigned)g.nz;
    int j = (int)(r % (unsigned)g.ny), i = (int)(r / (unsigned)g.ny);
    div[n] = divergence_at(g, u, v, w, dz_f, i + 1, j + 1, k + 1);
}

hipError_t run_div_field(const Grid& g, const double* u, const double* v, const double* w,
                         const double* dz_f, double* div, hipStream_t s) {
    long N = (long)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(div_field_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, g, u,
                       v, w, dz_f, div);
    return hipGetLastError();
}

// (nx, ny, nz) z-contiguous -> plane-major, divided by `scale`
__global__ __launch_bounds__(256) void to_planes_kernel(Grid g, const double* __restrict__ f,
                                                        double scale,
                                                        double* __restrict__ planes) {
    __shared__ double tile[kT][kT + 1];
    const int i = blockIdx.z;
    const int j0 = blockIdx.y * kT, k0 = blockIdx.x * kT;
    const int t = threadIdx.x;
    {
        const int kk = t % kT;
        for (int jj = t / kT; jj < kT; jj += 256 / kT) {
            int j = j0 + jj, k = k0 + kk;
            if (j < g.ny && k < g.nz) tile[jj][kk] = f[((long)i * g.ny + j) * g.nz + k] / scale;
        }
    }
    __syncthreads();
    {
        const int jj = t % kT;
        for (int kk = t / kT; kk < kT; kk += 256 / kT) {
            int j = j0 + jj, k = k0 + kk;
            if (j < g.ny && k < g.nz) planes[((long)k * g.nx + i) * g.ny + j] = tile[jj][kk];
        }
    }
}

hipError_t run_field_to_planes(const Grid& g, const double* f, double scale, double* planes,
                               hipStream_t s) {
    dim3 grid((g.nz + kT - 1) / kT, (g.ny + kT - 1) / kT, g.nx);
    hipLaunchKernelGGL(to_planes_kernel, grid, dim3(256), 0, s, g, f, scale, planes);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// per-wavenumber Thomas
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void poisson_thomas_kernel(PoissonTabs t,
                                                             double2* __restrict__ d,
                                                             double* __restrict__ cpw) {
    poisson_thomas_body(t, d, cpw, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

template <int NT>
__global__ __launch_bounds__(NT) void poisson_spike_kernel(PoissonTabs t,
                                                           double2* __restrict__ d) {
    __shared__ double lds[NT * 8];
    poisson_spike_body<NT>(t, d, (long)blockIdx.x, lds);
}

// Segments from which the partition solve runs 512-thread workgroups: at 64
// or more segments a 256-thread workgroup holds 4 modes or fewer, i.e. plane
// rows of 64 bytes or less.
#ifndef SLC_SPIKE_WIDE_SEGS
#define SLC_SPIKE_WIDE_SEGS 64
#endif

// The partition solve (32 B per unknown) where nz allows it, else the
// sequential sweep with c' through the cp scratch.
hipError_t run_poisson_thomas(const PoissonTabs& t, double* dhat, double* cp, hipStream_t s) {
    long M = (long)t.nx * t.nky;
    if (spike_ok(t.nz)) {
        double2* d = reinterpret_cast<double2*>(dhat);
        if (spike_segs(t.nz) >= SLC_SPIKE_WIDE_SEGS) {
            const long MB = spike_modes(t.nz, 512);
            hipLaunchKernelGGL(poisson_spike_kernel<512>, dim3((unsigned)((M + MB - 1) / MB)),
                               dim3(512), 0, s, t, d);
        } else {
            const long MB = spike_modes(t.nz);
            hipLaunchKernelGGL(poisson_spike_kernel<256>, dim3((unsigned)((M + MB - 1) / MB)),
                               dim3(256), 0, s, t, d);
        }
        return hipGetLastError();
    }
    hipLaunchKernelGGL(poisson_thomas_kernel, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, s,
                       t, reinterpret_cast<double2*>(dhat), cp);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// projection
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void project_planes_kernel(
    Grid g, double* __restrict__ u, double* __restrict__ v, double* __restrict__ w,
    const double* __restrict__ P, double inv_n, double dt, const double* __restrict__ dz_c,
    double* __restrict__ p) {
    __shared__ double t0[kT + 1][kT + 2];  // plane i   rows j0..j0+kT, k0..k0+kT
    __shared__ double t1[kT][kT + 1];      // plane i+1 rows j0..j0+kT-1
    project_planes_body(g, u, v, w, P, inv_n, dt, dz_c, p, blockIdx.z, blockIdx.y * kT,
                        blockIdx.x * kT, t0, t1);
}

hipError_t run_project_planes(const Grid& g, double* u, double* v, double* w,
                              const double* planes, double inv_n, double dt,
                              const double* dz_c, double* p, hipStream_t s) {
    dim3 grid((g.nz + kT - 1) / kT, (g.ny + kT - 1) / kT, g.nx);
    hipLaunchKernelGGL(project_planes_kernel, grid, dim3(256), 0, s, g, u, v, w, planes, inv_n, dt,
                       dz_c, p);
    return hipGetLastError();
}

// u -= dt grad p with a ghost-shaped p (projection.py:250-272)
__global__ void project_p_kernel(Grid g, double* __restrict__ u, double* __restrict__ v,
                                 double* __restrict__ w, const double* __restrict__ p, double dt,
                                 const double* __restrict__ dz_c) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (long)nx * ny * nz) return;
    const unsigned un = (unsigned)n;
    const int k = (int)(un % (unsigned)nz);
    const unsigned r = un / (unsigned)nz;
    const int j = (int)(r % (unsigned)ny), i = (int)(r / (unsigned)ny);
    const int I = i + 1, J = j + 1, K = k + 1;
    const long PSY = ny + 2, PSZ = nz + 2;
    const long pc = ((long)I * PSY + J) * PSZ + K;
    const double p0 = p[pc];
    long ui = ((long)I * (ny + 2) + J) * (nz + 2) + K;
    u[ui] = u[ui] - dt * ((p[pc + PSY * PSZ] - p0) / g.dx);
    long vi = ((long)I * (ny + 1) + J) * (nz + 2) + K;
    v[vi] = v[vi] - dt * ((p[pc + PSZ] - p0) / g.dy);
    if (K <= nz - 1) {
        long wi = ((long)I * (ny + 2) + J) * (nz + 1) + K;
        w[wi] = w[wi] - dt * ((p[pc + 1] - p0) / dz_c[K]);
    }
}

hipError_t run_project_p(const Grid& g, double* u, double* v, double* w, const double* p,
                         double dt, const double* dz_c, hipStream_t s) {
    long N = (long)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(project_p_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, g, u,
                       v, w, p, dt, dz_c);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// bulk velocity and the exact flux shift
// ---------------------------------------------------------------------------

int bulk_partials_needed(const Grid& g) { return bulk_blocks(g.nx, g.ny, g.nz); }

// shift 0: partial sums of u*cv; shift 1: u += corr (scal[0]) then partial sums
__global__ __launch_bounds__(kRedBlock) void bulk_partial_kernel(Grid g, double* __restrict__ u,
                                                                 const double* __restrict__ cv,
                                                                 const double* __restrict__ scal,
                                                                 int shift,
                                                                 double* __restrict__ partials) {
    __shared__ double red[kRedBlock];
    bulk_partial_body(g, u, cv, shift ? scal[0] : 0.0, shift, partials, blockIdx.x, gridDim.x,
                      red);
}

__global__ __launch_bounds__(kRedBlock) void bulk_finalize_kernel(
    const double* __restrict__ partials, int np, double total_volume, double U_bulk, double dt,
    int mode, double* __restrict__ scal) {
    __shared__ double red[kRedBlock];
    bulk_finalize_body(partials, np, total_volume, U_bulk, dt, mode, scal, red);
}

hipError_t run_bulk(const Grid& g, double* u, double* v, double* w, const double* cv,
                    double total_volume, double U_bulk, double dt, int shift, double* partials,
                    int max_partials, double* scal, hipStream_t s) {
    int np = bulk_partials_needed(g);
    if (np > max_partials) return hipErrorInvalidValue;
    hipLaunchKernelGGL(bulk_partial_kernel, dim3(np), dim3(kRedBlock), 0, s, g, u, cv, scal, 0,
                       partials);
    hipLaunchKernelGGL(bulk_finalize_kernel, dim3(1), dim3(kRedBlock), 0, s, partials, np,
                       total_volume, U_bulk, dt, 0, scal);
    if (!shift) return hipGetLastError();
    hipLaunchKernelGGL(bulk_partial_kernel, dim3(np), dim3(kRedBlock), 0, s, g, u, cv, scal, 1,
                       partials);
    hipLaunchKernelGGL(bulk_finalize_kernel, dim3(1), dim3(kRedBlock), 0, s, partials, np,
                       total_volume, U_bulk, dt, 1, scal);
    // u's ghosts follow the shifted interior (solver.py:720; v, w unchanged)
    (void)v;
    (void)w;
    return run_bc(g, u, nullptr, nullptr, s);
}

hipError_t run_bulk_velocity(const Grid& g, const double* u, const double* cv,
                             double total_volume, double* partials, int max_partials,
                             double* out, hipStream_t s) {
    int np = bulk_partials_needed(g);
    if (np > max_partials) return hipErrorInvalidValue;
    hipLaunchKernelGGL(bulk_partial_kernel, dim3(np), dim3(kRedBlock), 0, s, g,
                       const_cast<double*>(u), cv, out, 0, partials);
    // out[0..3] is used as scratch by mode 0; the bulk value lands in out[3]
    hipLaunchKernelGGL(bulk_finalize_kernel, dim3(1), dim3(kRedBlock), 0, s, partials, np,
                       total_volume, 0.0, 1.0, 0, out);
    return hipGetLastError();
}

}  // namespace slc
