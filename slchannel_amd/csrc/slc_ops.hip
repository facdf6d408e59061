// slc_ops.hip -- field operators of the SL-BDF2 step for gfx950:
//   ghost/boundary fill            solver.py:70-107, projection.py:180-195
//   theta-method z-diffusion       operators.py:382-630 (Thomas, shared factor)
//   divergence (+transpose)        utils.py:166-180
//   per-wavenumber Poisson Thomas  projection.py:133-247 (coefficients on the fly)
//   gradient projection            projection.py:250-272
//   bulk-flux constraint           solver.py:694-723, utils.py:147-163
//
// All of these are HBM-bound (well under one flop per byte); they are written
// for coalesced 64-lane access: z-contiguous fields are read/written along z,
// the spectral data lives plane-major (k, i, j) so rocFFT transforms contiguous
// x-y planes and the Thomas sweep over k reads consecutive modes per lane.
// Every pass is deterministic (fixed partition, fixed reduction trees).
#include "slc_internal.h"

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
// The reference assigns ghost slabs in a fixed order (x, then y, then z), so
// every ghost's final value is +-(one interior value).  Each thread writes one
// ghost entry from its interior source; interiors are only read, so there is
// no ordering hazard.  Entries are enumerated as
//   S1: whole x-ghost planes          (contiguous)
//   S2: y-ghost rows of interior x    (contiguous along z)
//   S3: the two z ends of interior columns.
// Field kinds: 0 u, 1 v, 2 w, 3 p.

struct GhostShape {
    int NI, NJ, NK;
    int nxg, xg0, xg1;  // x-ghost planes
    int nyg, yg0, yg1;  // y-ghost rows
    long S1, S2, S3;
};

__host__ __device__ inline GhostShape ghost_shape(int kind, int nx, int ny, int nz) {
    GhostShape s;
    s.NI = (kind == 0) ? nx + 1 : nx + 2;
    s.NJ = (kind == 1) ? ny + 1 : ny + 2;
    s.NK = (kind == 2) ? nz + 1 : nz + 2;
    if (kind == 0) { s.nxg = 1; s.xg0 = 0; s.xg1 = 0; }
    else { s.nxg = 2; s.xg0 = 0; s.xg1 = nx + 1; }
    if (kind == 1) { s.nyg = 1; s.yg0 = 0; s.yg1 = 0; }
    else { s.nyg = 2; s.yg0 = 0; s.yg1 = ny + 1; }
    s.S1 = (long)s.nxg * s.NJ * s.NK;
    s.S2 = (long)nx * s.nyg * s.NK;
    s.S3 = (long)nx * ny * 2;
    return s;
}

__device__ inline void ghost_value(int kind, const Grid& g, const double* F, const GhostShape& s,
                                   int I, int J, int K, double* out_val) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    int is = I, js = J, ks = K;
    double sg = 1.0;
    bool zero = false;
    if (I == 0) is = nx;
    else if (kind != 0 && I == nx + 1) is = 1;
    if (J == 0) js = ny;
    else if (kind != 1 && J == ny + 1) js = 1;
    if (kind == 2) {
        if (K == 0 || K == nz) zero = true;
    } else if (kind == 3) {
        if (K == 0) ks = 1;
        else if (K == nz + 1) ks = nz;
    } else {
        if (K == 0) { ks = 1; sg = -1.0; }
        else if (K == nz + 1) { ks = nz; sg = g.top_neumann ? 1.0 : -1.0; }
    }
    double v = zero ? 0.0 : F[((long)is * s.NJ + js) * s.NK + ks];
    *out_val = (sg < 0.0) ? -v : v;
}

__global__ void bc_kernel(Grid g, double* f0, double* f1, double* f2, int kind0) {
    const int kind = kind0 == 3 ? 3 : (int)blockIdx.y;
    double* F = kind0 == 3 ? f0 : (blockIdx.y == 0 ? f0 : (blockIdx.y == 1 ? f1 : f2));
    if (!F) return;
    GhostShape s = ghost_shape(kind, g.nx, g.ny, g.nz);
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int I, J, K;
    if (t < s.S1) {
        long pl = (long)s.NJ * s.NK;
        int p = (int)(t / pl);
        long r = t % pl;
        I = p == 0 ? s.xg0 : s.xg1;
        J = (int)(r / s.NK);
        K = (int)(r % s.NK);
    } else if (t < s.S1 + s.S2) {
        long t2 = t - s.S1;
        long row = t2 / s.NK;
        K = (int)(t2 % s.NK);
        I = 1 + (int)(row / s.nyg);
        J = (row % s.nyg) == 0 ? s.yg0 : s.yg1;
    } else if (t < s.S1 + s.S2 + s.S3) {
        long t3 = t - s.S1 - s.S2;
        long col = t3 >> 1;
        I = 1 + (int)(col / g.ny);
        J = 1 + (int)(col % g.ny);
        K = (t3 & 1) ? s.NK - 1 : 0;
    } else {
        return;
    }
    double v;
    ghost_value(kind, g, F, s, I, J, K, &v);
    F[((long)I * s.NJ + J) * s.NK + K] = v;
}

hipError_t run_bc(const Grid& g, double* u, double* v, double* w, hipStream_t st) {
    long nmax = 0;
    for (int k = 0; k < 3; ++k) {
        GhostShape s = ghost_shape(k, g.nx, g.ny, g.nz);
        long n = s.S1 + s.S2 + s.S3;
        if (n > nmax) nmax = n;
    }
    hipLaunchKernelGGL(bc_kernel, dim3((unsigned)((nmax + 255) / 256), 3), dim3(256), 0, st, g, u,
                       v, w, 0);
    return hipGetLastError();
}

hipError_t run_bc_p(const Grid& g, double* p, hipStream_t st) {
    GhostShape s = ghost_shape(3, g.nx, g.ny, g.nz);
    long n = s.S1 + s.S2 + s.S3;
    hipLaunchKernelGGL(bc_kernel, dim3((unsigned)((n + 255) / 256), 1), dim3(256), 0, st, g, p,
                       nullptr, nullptr, 3);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// implicit z-diffusion: (I - theta dt nu Dzz) x = d + (1-theta) dt nu Dzz d
// ---------------------------------------------------------------------------
// The matrix is the same for every (i, j) column, so it is factored once per
// call by one thread (tab = [a | c' | 1/den], n entries each) and every column
// only runs forward/back substitution.  Columns are staged through LDS so the
// HBM traffic is one coalesced read and one coalesced write per unknown.

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

constexpr int kZColsMax = 16;
constexpr int kZLdsBytes = 48 * 1024;  // dynamic LDS budget per workgroup

template <int C>
__global__ __launch_bounds__(256) void zsolve_kernel(Grid g, double* __restrict__ F,
                                                     const double* __restrict__ tab, double beta,
                                                     const double* __restrict__ dz_c,
                                                     const double* __restrict__ dz_f,
                                                     int kZCols) {
    extern __shared__ double lds[];
    const int SZ = (int)Comp<C>::SZ(g);
    const long SY = Comp<C>::SY(g);
    const int ld = SZ + ((SZ & 1) ? 0 : 1);  // odd stride: conflict-free column walk
    const int n = (C == 2) ? g.nz - 1 : g.nz;
    const long ncols = (long)g.nx * g.ny;
    const long col0 = (long)blockIdx.x * kZCols;
    const int tid = threadIdx.x;

    // column base addresses (one 32-bit division per column, not per element)
    __shared__ long cbase[kZColsMax];
    if (tid < kZCols) {
        const long gc = col0 + tid;
        long a = -1;
        if (gc < ncols) {
            const unsigned ug = (unsigned)gc;
            const int I = 1 + (int)(ug / (unsigned)g.ny), J = 1 + (int)(ug % (unsigned)g.ny);
            a = ((long)I * SY + J) * SZ;
        }
        cbase[tid] = a;
    }
    __syncthreads();
    // coalesced load of kZCols whole columns (ghost ends included)
    for (int e = tid; e < kZCols * SZ; e += blockDim.x) {
        const int c = e / SZ, k = e - (e / SZ) * SZ;
        const long a = cbase[c];
        if (a >= 0) lds[c * ld + k] = F[a + k];
    }
    __syncthreads();

    if (tid < kZCols && col0 + tid < ncols) {
        double* s = lds + tid * ld;
        if (beta != 0.0) {
            // explicit part, operators.py:446-466 (u, v) and 609-623 (w)
            if (C < 2) {
                double lo = s[0], mid = s[1];
                for (int k = 0; k < n; ++k) {
                    double hi = s[k + 2];
                    double dzl = dz_c[k], dzr = dz_c[k + 1], dzc = dz_f[k];
                    double d2;
                    if (k == 0) {
                        d2 = ((hi - mid) / dzr - 2.0 * mid / dzl) / dzc;
                    } else if (k == n - 1) {
                        if (g.top_neumann) d2 = (-(mid - lo) / dzl) / dzc;
                        else d2 = (-2.0 * mid / dzr - (mid - lo) / dzl) / dzc;
                    } else {
                        d2 = ((hi - mid) / dzr - (mid - lo) / dzl) / dzc;
                    }
                    s[k + 1] = mid + beta * d2;
                    lo = mid;
                    mid = hi;
                }
            } else {
                // the reference's w stencil is centred on face m (not m+1)
                double a = 0.0, b = s[0], c = s[1];
                for (int m = 0; m < n; ++m) {
                    double dzl = dz_f[m], dzr = dz_f[m + 1], dzc = dz_c[m + 1];
                    double d2;
                    if (m == 0) d2 = ((c - b) / dzr - b / dzl) / dzc;
                    else if (m == n - 1) d2 = (-b / dzr - (b - a) / dzl) / dzc;
                    else d2 = ((c - b) / dzr - (b - a) / dzl) / dzc;
                    double nxt = s[m + 2];
                    s[m + 1] = c + beta * d2;
                    a = b;
                    b = c;
                    c = nxt;
                }
            }
        }
        const double* A = tab;
        const double* CP = tab + n;
        const double* INV = tab + 2 * n;
        double dp = s[1] * INV[0];
        s[1] = dp;
        for (int k = 1; k < n; ++k) {
            dp = (s[k + 1] - A[k] * dp) * INV[k];
            s[k + 1] = dp;
        }
        double x = dp;
        for (int k = n - 2; k >= 0; --k) {
            x = s[k + 1] - CP[k] * x;
            s[k + 1] = x;
        }
    }
    __syncthreads();

    for (int e = tid; e < kZCols * SZ; e += blockDim.x) {
        const int c = e / SZ, k = e - (e / SZ) * SZ;
        const long a = cbase[c];
        if (a >= 0 && k >= 1 && k <= n) F[a + k] = lds[c * ld + k];
    }
}

hipError_t run_zsolve(const Grid& g, int comp, double* f, double dt, double theta,
                      const double* dz_c, const double* dz_f, double* tab, bool refactor,
                      hipStream_t s) {
    if (refactor)
        hipLaunchKernelGGL(zfactor_kernel, dim3(1), dim3(64), 0, s, g, comp, dt, theta, dz_c,
                           dz_f, tab);
    double beta = theta < 1.0 ? (1.0 - theta) * dt * g.nu : 0.0;
    long ncols = (long)g.nx * g.ny;
    int SZ = comp == 2 ? g.nz + 1 : g.nz + 2;
    int ld = SZ + ((SZ & 1) ? 0 : 1);
    int cols = kZLdsBytes / (ld * (int)sizeof(double));
    if (cols > kZColsMax) cols = kZColsMax;
    if (cols < 1) return hipErrorInvalidValue;  // nz beyond ~6000: column does not fit LDS
    dim3 grid((unsigned)((ncols + cols - 1) / cols));
    size_t shm = (size_t)cols * ld * sizeof(double);
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

constexpr int kT = 32;  // transpose tile edge

__device__ inline double divergence_at(const Grid& g, const double* __restrict__ u,
                                       const double* __restrict__ v,
                                       const double* __restrict__ w,
                                       const double* __restrict__ dz_f, int I, int J, int K) {
    const long uSY = g.ny + 2, uSZ = g.nz + 2;
    const long vSY = g.ny + 1, vSZ = g.nz + 2;
    const long wSY = g.ny + 2, wSZ = g.nz + 1;
    double du = (u[((long)I * uSY + J) * uSZ + K] - u[((long)(I - 1) * uSY + J) * uSZ + K]) / g.dx;
    double dv = (v[((long)I * vSY + J) * vSZ + K] - v[((long)I * vSY + (J - 1)) * vSZ + K]) / g.dy;
    long wb = ((long)I * wSY + J) * wSZ;
    double dw = (w[wb + K] - w[wb + K - 1]) / dz_f[K - 1];
    return du + dv + dw;
}

__global__ __launch_bounds__(256) void div_planes_kernel(Grid g, const double* __restrict__ u,
                                                         const double* __restrict__ v,
                                                         const double* __restrict__ w,
                                                         const double* __restrict__ dz_f,
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
            if (j < g.ny && k < g.nz) {
                double d = divergence_at(g, u, v, w, dz_f, i + 1, j + 1, k + 1);
                tile[jj][kk] = d / scale;
            }
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
// per-wavenumber Thomas (projection.py:199-247 with the matrices of 47-114)
// ---------------------------------------------------------------------------
// dhat: (nz, nx, nky) complex, plane-major; one lane per mode walks k, so a
// wavefront reads 64 consecutive modes (1 KiB) per plane.  c' is kept in a
// real scratch of the same (nz, modes) shape.

__global__ __launch_bounds__(256) void poisson_thomas_kernel(PoissonTabs t,
                                                             double2* __restrict__ d,
                                                             double* __restrict__ cpw) {
    const long M = (long)t.nx * t.nky;
    long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    const int nz = t.nz;
    const int i = (int)(m / t.nky), jc = (int)(m % t.nky);
    const double khsq = t.kxsq[i] + t.kysq[jc];
    // row 0 (Neumann bottom; identity row for the (0,0) mode)
    double b = -(t.cs[0] + khsq);
    b += t.cl[0];
    if (nz == 1) b += t.cr[0];
    double c = nz > 1 ? t.cr[0] : 0.0;
    if (m == 0) {
        b = 1.0;
        c = 0.0;
    }
    double cp = c / b;
    double2 dk = d[m];
    double dr = dk.x / b, di = dk.y / b;
    d[m] = make_double2(dr, di);
    cpw[m] = cp;
    for (int k = 1; k < nz; ++k) {
        const double a = t.cl[k];
        double bk = -(t.cs[k] + khsq);
        double ck = 0.0;
        if (k == nz - 1) bk += t.cr[nz - 1];
        else ck = t.cr[k];
        double den = bk - a * cp;
        if (k < nz - 1) cp = ck / den;
        long idx = (long)k * M + m;
        dk = d[idx];
        dr = (dk.x - a * dr) / den;
        di = (dk.y - a * di) / den;
        d[idx] = make_double2(dr, di);
        cpw[idx] = cp;
    }
    double xr = dr, xi = di;
    for (int k = nz - 2; k >= 0; --k) {
        long idx = (long)k * M + m;
        double cpk = cpw[idx];
        dk = d[idx];
        xr = dk.x - cpk * xr;
        xi = dk.y - cpk * xi;
        d[idx] = make_double2(xr, xi);
    }
}

hipError_t run_poisson_thomas(const PoissonTabs& t, double* dhat, double* cp, hipStream_t s) {
    long M = (long)t.nx * t.nky;
    hipLaunchKernelGGL(poisson_thomas_kernel, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, s,
                       t, reinterpret_cast<double2*>(dhat), cp);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// projection u -= dt grad p from the plane-major (unnormalised) inverse FFT
// ---------------------------------------------------------------------------
// Also writes the interior of the ghost-shaped pressure p = P / (nx ny).

__global__ __launch_bounds__(256) void project_planes_kernel(
    Grid g, double* __restrict__ u, double* __restrict__ v, double* __restrict__ w,
    const double* __restrict__ P, double inv_n, double dt, const double* __restrict__ dz_c,
    double* __restrict__ p) {
    __shared__ double t0[kT + 1][kT + 2];  // plane i   rows j0..j0+kT, k0..k0+kT
    __shared__ double t1[kT][kT + 1];      // plane i+1 rows j0..j0+kT-1
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int i = blockIdx.z;
    const int i1 = (i + 1 == nx) ? 0 : i + 1;
    const int j0 = blockIdx.y * kT, k0 = blockIdx.x * kT;
    const int t = threadIdx.x;
    // load (j fastest: plane rows are contiguous in j)
    for (int e = t; e < (kT + 1) * (kT + 1); e += 256) {
        int jj = e % (kT + 1), kk = e / (kT + 1);
        int j = j0 + jj, k = k0 + kk;
        if (k < nz && j <= ny && jj <= kT) {
            int jw = (j == ny) ? 0 : j;
            t0[jj][kk] = P[((long)k * nx + i) * ny + jw];
        }
    }
    for (int e = t; e < kT * kT; e += 256) {
        int jj = e % kT, kk = e / kT;
        int j = j0 + jj, k = k0 + kk;
        if (k < nz && j < ny) t1[jj][kk] = P[((long)k * nx + i1) * ny + j];
    }
    __syncthreads();
    // update (k fastest: fields are z-contiguous)
    const int I = i + 1;
    const int kk = t % kT;
    for (int jj = t / kT; jj < kT; jj += 256 / kT) {
        int j = j0 + jj, k = k0 + kk;
        if (j >= ny || k >= nz) continue;
        const int J = j + 1, K = k + 1;
        double pc = t0[jj][kk] * inv_n;
        double pxp = t1[jj][kk] * inv_n;
        double pyp = t0[jj + 1][kk] * inv_n;
        if (u) {
            long ui = ((long)I * (ny + 2) + J) * (nz + 2) + K;
            u[ui] = u[ui] - dt * ((pxp - pc) / g.dx);
            long vi = ((long)I * (ny + 1) + J) * (nz + 2) + K;
            v[vi] = v[vi] - dt * ((pyp - pc) / g.dy);
        }
        if (u && k < nz - 1) {
            double pzp = t0[jj][kk + 1] * inv_n;
            long wi = ((long)I * (ny + 2) + J) * (nz + 1) + K;
            w[wi] = w[wi] - dt * ((pzp - pc) / dz_c[K]);
        }
        if (p) p[((long)I * (ny + 2) + J) * (nz + 2) + K] = pc;
    }
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
// sum_{ijk} u_ijk * (dx dy dz_f[k]) / V, reduced over a fixed partition and a
// fixed tree (deterministic).  scal: [0] correction, [1] forcing, [2] u_bulk
// after the shift, [3] u_bulk before it.

constexpr int kRedBlock = 256;

int bulk_partials_needed(const Grid& g) {
    long N = (long)g.nx * g.ny * g.nz;
    long nb = (N + kRedBlock * 8 - 1) / (kRedBlock * 8);
    return (int)(nb < 2048 ? nb : 2048);
}

__device__ inline double block_sum(double v) {
    __shared__ double red[kRedBlock];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kRedBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    return red[0];
}

// mode 0: partial sums of u*cv; mode 1: u += corr (scal[0]) then partial sums
__global__ __launch_bounds__(kRedBlock) void bulk_partial_kernel(Grid g, double* __restrict__ u,
                                                                 const double* __restrict__ cv,
                                                                 const double* __restrict__ scal,
                                                                 int shift,
                                                                 double* __restrict__ partials) {
    const long N = (long)g.nx * g.ny * g.nz;
    const long nb = gridDim.x;
    const long chunk = (N + nb - 1) / nb;
    const long beg = blockIdx.x * chunk;
    const long end = beg + chunk < N ? beg + chunk : N;
    const double corr = shift ? scal[0] : 0.0;
    double acc = 0.0;
    for (long n = beg + threadIdx.x; n < end; n += kRedBlock) {
        const unsigned un = (unsigned)n;
        int k = (int)(un % (unsigned)g.nz);
        unsigned r = un / (unsigned)g.nz;
        int j = (int)(r % (unsigned)g.ny), i = (int)(r / (unsigned)g.ny);
        long ui = ((long)(i + 1) * (g.ny + 2) + (j + 1)) * (g.nz + 2) + (k + 1);
        double val = u[ui];
        if (shift) {
            val = val + corr;
            u[ui] = val;
        }
        acc += val * cv[k];
    }
    double s = block_sum(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// mode 0: scal[3] = ub*, scal[0] = U_b - ub*, scal[1] = scal[0] / dt
// mode 1: scal[2] = ub
__global__ __launch_bounds__(kRedBlock) void bulk_finalize_kernel(
    const double* __restrict__ partials, int np, double total_volume, double U_bulk, double dt,
    int mode, double* __restrict__ scal) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < np; b += kRedBlock) acc += partials[b];
    double s = block_sum(acc);
    if (threadIdx.x == 0) {
        double ub = s / total_volume;
        if (mode == 0) {
            double corr = U_bulk - ub;
            scal[3] = ub;
            scal[0] = corr;
            scal[1] = corr / dt;
        } else {
            scal[2] = ub;
        }
    }
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
