// slc_init.hip -- seeded solenoidal initial fields on the device
// (slc_init_solenoidal).
//
// The perturbation is the curl of a random vector potential
//   A_c(x, y, z) = sum_b Re[a^c_b e^{i(kx_m x + ky_n y)}] G_l(z),
// b = (m, n) a wavenumber bin, l = 1..L the wall-normal modes.  The host draws
// the amplitudes (seeded) and hands over, per velocity component, the
// coefficients of G_l and dG_l/dz in each bin (alpha, beta: the curl taken on
// the amplitudes); this file evaluates the velocity at every staggered
// interior position by direct summation, factored by axis:
//   coef_c[b][k] = sum_l alpha_c[b][l] G_l(z_k) + beta_c[b][l] G'_l(z_k)
//   T_c[m][j][k] = sum_n Ey_c[n][j] coef_c[m W + n][k]
//   f_c(i, j, k) = sum_m Re(Ex_c[m][i] T_c[m][j][k])
// with Ex_c[m][i] = e^{i kx_m x_i}, Ey_c[n][j] = e^{i ky_n y_j} at the
// component's staggered positions (tables from the host).  A last pass scales
// the three components to the requested u'rms (sum of squares over the
// interiors, reduced in a fixed order) and adds the Poiseuille base profile to
// u.  Four launches and no host synchronisation, whatever the number of modes:
// an admission in a serving loop costs a few launches instead of one torch
// operation per step of the synthesis.
#include "slc_plan.h"

using namespace slc;

namespace {

constexpr int kIB = 256;           // threads per workgroup
constexpr int kInitBlocks = 1024;  // workgroups per component in the two field passes

__device__ inline double2 cfma(double2 a, double2 b, double2 acc) {
    return make_double2(fma(a.x, b.x, fma(-a.y, b.y, acc.x)), fma(a.x, b.y, fma(a.y, b.x, acc.y)));
}

// ab: five (W^2, L) complex blocks: alpha_u, beta_u, alpha_v, beta_v, alpha_w.
// zt: G at the centres (L, nz), G' at the centres (L, nz), G at the interior
// faces (L, nz-1), the base profile (nz).  coef: (3, W^2, nz) complex.
__global__ __launch_bounds__(kIB) void init_coef_kernel(int W, int L, int nz,
                                                        const double2* __restrict__ ab,
                                                        const double* __restrict__ zt,
                                                        double2* __restrict__ coef) {
    const int c = blockIdx.y;
    const long nzc = c == 2 ? nz - 1 : nz;
    const long WW = (long)W * W;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= WW * nzc) return;
    const long b = t / nzc, k = t - b * nzc;
    const double* g = c == 2 ? zt + 2L * L * nz : zt;
    const double* dg = zt + (long)L * nz;
    const long gs = c == 2 ? nz - 1 : nz;  // row stride of g
    const double2* al = ab + (2 * c) * WW * L + b * L;
    const double2* be = c < 2 ? ab + (2 * c + 1) * WW * L + b * L : al;  // w: no G' term
    double2 acc = make_double2(0.0, 0.0);
    for (int l = 0; l < L; ++l) {
        const double gv = g[l * gs + k];
        acc.x = fma(al[l].x, gv, acc.x);
        acc.y = fma(al[l].y, gv, acc.y);
        if (c < 2) {
            const double dv = dg[l * gs + k];
            acc.x = fma(be[l].x, dv, acc.x);
            acc.y = fma(be[l].y, dv, acc.y);
        }
    }
    coef[(c * WW + b) * nz + k] = acc;
}

// phase: Ex_u (W, nx), Ey_u (W, ny), Ex_v, Ey_v, Ex_w, Ey_w (complex).
// T: (3, W, ny, nz) complex.
__global__ __launch_bounds__(kIB) void init_y_kernel(int W, int nx, int ny, int nz,
                                                     const double2* __restrict__ coef,
                                                     const double2* __restrict__ phase,
                                                     double2* __restrict__ T) {
    const int c = blockIdx.y;
    const long nzc = c == 2 ? nz - 1 : nz;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)W * ny * nzc) return;
    const long m = t / (ny * nzc), r = t - m * ny * nzc;
    const long j = r / nzc, k = r - j * nzc;
    const long WW = (long)W * W;
    const double2* Ey = phase + (long)c * W * (nx + ny) + (long)W * nx;
    const double2* cf = coef + (c * WW + m * W) * nz + k;
    double2 acc = make_double2(0.0, 0.0);
    for (int n = 0; n < W; ++n) acc = cfma(Ey[(long)n * ny + j], cf[(long)n * nz], acc);
    T[((c * W + m) * ny + j) * nz + k] = acc;
}

__device__ inline long field_index(int c, int nx, int ny, int nz, long i, long j, long k) {
    const long SY = c == 1 ? ny + 1 : ny + 2;
    const long SZ = c == 2 ? nz + 1 : nz + 2;
    return ((i + 1) * SY + (j + 1)) * SZ + (k + 1);
}

__device__ inline double block_sum_init(double v, double* red) {
    __syncthreads();
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kIB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    return red[0];
}

// f_c at the interior points (written unscaled) and per-workgroup sums of
// squares: part[c * kInitBlocks + blockIdx.x].  A work item is a run of kIX
// consecutive x positions at one (j, k): T_c[m][j][k] does not depend on x, so
// it is read once per run (z-contiguous across the lanes) and applied to the
// run's kIX accumulators, each summed over m in the same order as a single
// point would be (reading T once per point made the pass bound by T's
// traffic: 296 ms at 1024x1024x512, 4.3 ms at 256^3).
constexpr int kIX = 16;

__global__ __launch_bounds__(kIB) void init_x_kernel(int W, int nx, int ny, int nz,
                                                     const double2* __restrict__ T,
                                                     const double2* __restrict__ phase,
                                                     double* __restrict__ u,
                                                     double* __restrict__ v,
                                                     double* __restrict__ w,
                                                     double* __restrict__ part) {
    __shared__ double red[kIB];
    const int c = blockIdx.y;
    const long nzc = c == 2 ? nz - 1 : nz;
    const long nrun = (nx + kIX - 1) / kIX;
    const long N = nrun * ny * nzc;
    const double2* Ex = phase + (long)c * W * (nx + ny);
    double* F = c == 0 ? u : (c == 1 ? v : w);
    double sq = 0.0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < N;
         t += (long)gridDim.x * blockDim.x) {
        const long ir = t / (ny * nzc), r = t - ir * ny * nzc;
        const long j = r / nzc, k = r - j * nzc;
        const long i0 = ir * kIX;
        const int ni = (int)(nx - i0 < kIX ? nx - i0 : kIX);
        const double2* Tc = T + (c * (long)W * ny + j) * nz + k;
        double val[kIX];
#pragma unroll
        for (int q = 0; q < kIX; ++q) val[q] = 0.0;
        for (int m = 0; m < W; ++m) {
            const double2 tv = Tc[(long)m * ny * nz];
            const double2* E = Ex + (long)m * nx + i0;
#pragma unroll
            for (int q = 0; q < kIX; ++q) {
                if (q < ni) {
                    const double2 e = E[q];
                    val[q] = fma(e.x, tv.x, fma(-e.y, tv.y, val[q]));
                }
            }
        }
#pragma unroll
        for (int q = 0; q < kIX; ++q) {
            if (q < ni) {
                F[field_index(c, nx, ny, nz, i0 + q, j, k)] = val[q];
                sq = fma(val[q], val[q], sq);
            }
        }
    }
    const double s = block_sum_init(sq, red);
    if (threadIdx.x == 0) part[c * kInitBlocks + blockIdx.x] = s;
}

// f_c *= urms / rms (rms over the three interiors; every workgroup reduces the
// partials in the same order), plus the base profile for u.
__global__ __launch_bounds__(kIB) void init_scale_kernel(int nx, int ny, int nz, double urms,
                                                         const double* __restrict__ base,
                                                         const double* __restrict__ part,
                                                         double* __restrict__ u,
                                                         double* __restrict__ v,
                                                         double* __restrict__ w) {
    __shared__ double red[kIB];
    double acc = 0.0;
    for (int q = threadIdx.x; q < 3 * kInitBlocks; q += kIB) acc += part[q];
    const double total = block_sum_init(acc, red);
    const double count = (double)nx * ny * (2.0 * nz + (nz - 1));
    const double ss = total / count;
    const double scale = ss > 0.0 ? urms / sqrt(ss) : 0.0;
    const int c = blockIdx.y;
    const long nzc = c == 2 ? nz - 1 : nz;
    const long N = (long)nx * ny * nzc;
    double* F = c == 0 ? u : (c == 1 ? v : w);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < N;
         t += (long)gridDim.x * blockDim.x) {
        const long i = t / (ny * nzc), r = t - i * ny * nzc;
        const long j = r / nzc, k = r - j * nzc;
        const long idx = field_index(c, nx, ny, nz, i, j, k);
        const double f = scale * F[idx];
        F[idx] = c == 0 ? f + base[k] : f;
    }
}

}  // namespace

extern "C" int64_t slc_init_solenoidal_work(int nx, int ny, int nz, int W) {
    if (nx < 1 || ny < 1 || nz < 2 || W < 1) return -1;
    return 16L * 3 * W * W * nz + 16L * 3 * W * ny * nz + 8L * 3 * kInitBlocks;
}

extern "C" int slc_init_solenoidal(slc_plan* p, int W, int L, const double* ab,
                                   const double* zt, const double* phase, double urms,
                                   double* work, int64_t work_bytes, double* u, double* v,
                                   double* w, void* stream) {
    if (!p) return fail(SLC_EINVAL, "null plan");
    if (!ab || !zt || !phase || !work || !u || !v || !w) return fail(SLC_EINVAL, "null array");
    if (W < 1 || W % 2 == 0 || L < 1) return fail(SLC_EINVAL, "bins W=%d (odd) and L=%d", W, L);
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const int64_t need = slc_init_solenoidal_work(nx, ny, nz, W);
    if (need < 0) return fail(SLC_EINVAL, "grid too small");
    if (work_bytes < need)
        return fail(SLC_EINVAL, "work buffer of %lld bytes, %lld needed", (long long)work_bytes,
                    (long long)need);
    DeviceGuard dg(p->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    double2* coef = reinterpret_cast<double2*>(work);
    double2* T = coef + 3L * W * W * nz;
    double* part = reinterpret_cast<double*>(T + 3L * W * ny * nz);
    const double2* ab2 = reinterpret_cast<const double2*>(ab);
    const double2* ph2 = reinterpret_cast<const double2*>(phase);
    auto blocks = [](long n) { return (unsigned)((n + kIB - 1) / kIB); };
    hipLaunchKernelGGL(init_coef_kernel, dim3(blocks((long)W * W * nz), 3), dim3(kIB), 0, s, W,
                       L, nz, ab2, zt, coef);
    HIP_OK(hipGetLastError());
    hipLaunchKernelGGL(init_y_kernel, dim3(blocks((long)W * ny * nz), 3), dim3(kIB), 0, s, W, nx,
                       ny, nz, coef, ph2, T);
    HIP_OK(hipGetLastError());
    hipLaunchKernelGGL(init_x_kernel, dim3(kInitBlocks, 3), dim3(kIB), 0, s, W, nx, ny, nz, T,
                       ph2, u, v, w, part);
    HIP_OK(hipGetLastError());
    const double* base = zt + 2L * L * nz + (long)L * (nz - 1);
    hipLaunchKernelGGL(init_scale_kernel, dim3(kInitBlocks, 3), dim3(kIB), 0, s, nx, ny, nz, urms,
                       base, part, u, v, w);
    HIP_OK(hipGetLastError());
    return SLC_OK;
}

// ---------------------------------------------------------------------------
// trilinear interpolation onto tensor-product nodes (restart onto another grid)
// ---------------------------------------------------------------------------

namespace {

// out[a][b][c] = sum over p, q, r in {0, 1} of ((wx_p wy_q) wz_r) src[ix+p][iy+q][iz+r]
// with wx_0 = 1 - tx, wx_1 = tx (likewise y, z), accumulated from zero in
// (p, q, r) order: the operations of initflow.py:31-44.  An index whose
// upper node is outside the source writes NaN.
__global__ __launch_bounds__(kIB) void interp_tri_kernel(
    const double* __restrict__ src, int sx, int sy, int sz, const int* __restrict__ ix,
    const double* __restrict__ tx, int nxq, const int* __restrict__ iy,
    const double* __restrict__ ty, int nyq, const int* __restrict__ iz,
    const double* __restrict__ tz, int nzq, double* __restrict__ out) {
    const long N = (long)nxq * nyq * nzq;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const int c = (int)(t % nzq);
    const long r = t / nzq;
    const int b = (int)(r % nyq), a = (int)(r / nyq);
    const int i0 = ix[a], j0 = iy[b], k0 = iz[c];
    if (i0 < 0 || i0 + 1 >= sx || j0 < 0 || j0 + 1 >= sy || k0 < 0 || k0 + 1 >= sz) {
        out[t] = __builtin_nan("");
        return;
    }
    const double wx[2] = {1.0 - tx[a], tx[a]};
    const double wy[2] = {1.0 - ty[b], ty[b]};
    const double wz[2] = {1.0 - tz[c], tz[c]};
    double acc = 0.0;
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int rr = 0; rr < 2; ++rr)
                acc = acc + ((wx[p] * wy[q]) * wz[rr]) *
                                src[((long)(i0 + p) * sy + (j0 + q)) * sz + (k0 + rr)];
    out[t] = acc;
}

}  // namespace

extern "C" int slc_interp_trilinear(int device, const double* src, int sx, int sy, int sz,
                                    const int* ix, const double* tx, int nxq, const int* iy,
                                    const double* ty, int nyq, const int* iz, const double* tz,
                                    int nzq, double* out, void* stream) {
    if (!src || !ix || !tx || !iy || !ty || !iz || !tz || !out)
        return fail(SLC_EINVAL, "null array");
    if (sx < 2 || sy < 2 || sz < 2 || nxq < 1 || nyq < 1 || nzq < 1)
        return fail(SLC_EINVAL, "source %dx%dx%d (each >= 2), query %dx%dx%d (each >= 1)", sx, sy,
                    sz, nxq, nyq, nzq);
    DeviceGuard dg(device);
    const long N = (long)nxq * nyq * nzq;
    hipLaunchKernelGGL(interp_tri_kernel, dim3((unsigned)((N + kIB - 1) / kIB)), dim3(kIB), 0,
                       static_cast<hipStream_t>(stream), src, sx, sy, sz, ix, tx, nxq, iy, ty, nyq,
                       iz, tz, nzq, out);
    HIP_OK(hipGetLastError());
    return SLC_OK;
}
