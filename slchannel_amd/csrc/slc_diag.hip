// slc_diag.hip -- run-loop diagnostics on the device (slc_diagnostics).
//
// One pass over the cells gives what the reference's run loop prints every
// n_out steps and checks every n_save steps (solver.py:1355-1479):
//   max|div|  compute_divergence (utils.py:166-180), element for element;
//   dti       compute_cfl_fused (operators.py:839-940): per cell the x, y
//             and z inverse time-scales with the same 4-point averages and
//             operation order, maximum over cells;
//   u_tau     compute_u_tau (utils.py:116-146): plane means of u at the
//             first and last interior levels over the whole ghost-shaped
//             x-y extent, as the reference takes them;
//   non-finite entries of u, v, w, p (the NaN/Inf guard, solver.py:1436-1479).
// A second one-workgroup kernel reduces the per-workgroup partials in a fixed
// order and writes the results to device memory: no host synchronisation.
// Maxima propagate NaN like torch.max.
#include <cfloat>

#include "slc_plan.h"

using namespace slc;

namespace {

constexpr int kDiagBlock = 256;

__device__ inline double nanmax(double a, double b) {
    return (a != a || a > b) ? a : b;  // NaN wins, as torch.max
}

__device__ inline bool nonfinite(double x) { return !(fabs(x) <= DBL_MAX); }

struct DiagPart {
    double div, tix, tiy, tiz, bad;
};

__global__ __launch_bounds__(kDiagBlock) void diag_partial_kernel(
    Grid g, const double* __restrict__ u, const double* __restrict__ v,
    const double* __restrict__ w, const double* __restrict__ p, const double* __restrict__ dzf,
    const double* __restrict__ dzc, double* __restrict__ part) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long su = (long)(ny + 2) * (nz + 2), sv = (long)(ny + 1) * (nz + 2),
               sw = (long)(ny + 2) * (nz + 1);
    // plans keep every array below 2^31 elements: 32-bit cell decode
    const unsigned ncell = (unsigned)nx * ny * nz;
    const double dxi = 1.0 / g.dx, dyi = 1.0 / g.dy;
    double mdiv = 0.0, mx = 0.0, my = 0.0, mz = 0.0, bad = 0.0;
    const unsigned stride = gridDim.x * blockDim.x;
    for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < ncell; t += stride) {
        const int k = (int)(t % (unsigned)nz) + 1;
        const unsigned r = t / (unsigned)nz;
        const int j = (int)(r % (unsigned)ny) + 1, i = (int)(r / (unsigned)ny) + 1;
        // index helpers: U(i,j,k), V(i,j,k), W(i,j,k)
        const long uo = i * su + (long)j * (nz + 2) + k;
        const long vo = i * sv + (long)j * (nz + 2) + k;
        const long wo = i * sw + (long)j * (nz + 1) + k;
        const double dzfi = 1.0 / dzf[k - 1];
        // divergence (utils.py:175-179)
        const double du = (u[uo] - u[uo - su]) / g.dx;
        const double dv = (v[vo] - v[vo - (nz + 2)]) / g.dy;
        const double dw = (w[wo] - w[wo - 1]) / dzf[k - 1];
        mdiv = nanmax(mdiv, fabs(du + dv + dw));
        // x family (operators.py:870-886)
        {
            const double ux = fabs(u[uo]);
            const double vx = 0.25 * fabs(v[vo] + v[vo - (nz + 2)] + v[vo + sv] +
                                          v[vo + sv - (nz + 2)]);
            const double wx = 0.25 * fabs(w[wo] + w[wo - 1] + w[wo + sw] + w[wo + sw - 1]);
            mx = nanmax(mx, ux * dxi + vx * dyi + wx * dzfi);
        }
        // y family (operators.py:891-907)
        {
            const double uy = 0.25 * fabs(u[uo] + u[uo + (nz + 2)] + u[uo - su + (nz + 2)] +
                                          u[uo - su]);
            const double vy = fabs(v[vo]);
            const double wy = 0.25 * fabs(w[wo] + w[wo + (nz + 1)] + w[wo + (nz + 1) - 1] +
                                          w[wo - 1]);
            my = nanmax(my, uy * dxi + vy * dyi + wy * dzfi);
        }
        // z family at the interior faces k = 1..nz-1 (operators.py:912-928)
        if (k < nz) {
            const double uz = 0.25 * fabs(u[uo] + u[uo - su] + u[uo - su + 1] + u[uo + 1]);
            const double vz = 0.25 * fabs(v[vo] + v[vo - (nz + 2)] + v[vo - (nz + 2) + 1] +
                                          v[vo + 1]);
            const double wz = fabs(w[wo]);
            mz = nanmax(mz, uz * dxi + vz * dyi + wz * (1.0 / dzc[k]));
        }
    }
    // non-finite entries anywhere in the four arrays (ghosts included)
    const long n[4] = {(long)(nx + 1) * su, (long)(nx + 2) * sv, (long)(nx + 2) * sw,
                       (long)(nx + 2) * (ny + 2) * (nz + 2)};
    const double* a[4] = {u, v, w, p};
    for (int c = 0; c < 4; ++c) {
        if (!a[c]) continue;
        for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n[c]; t += (long)stride)
            bad += nonfinite(a[c][t]) ? 1.0 : 0.0;
    }
    __shared__ DiagPart red[kDiagBlock];
    red[threadIdx.x] = DiagPart{mdiv, mx, my, mz, bad};
    __syncthreads();
    for (int s = kDiagBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            DiagPart& x = red[threadIdx.x];
            const DiagPart& y = red[threadIdx.x + s];
            x.div = nanmax(x.div, y.div);
            x.tix = nanmax(x.tix, y.tix);
            x.tiy = nanmax(x.tiy, y.tiy);
            x.tiz = nanmax(x.tiz, y.tiz);
            x.bad += y.bad;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* o = part + 5 * (long)blockIdx.x;
        o[0] = red[0].div;
        o[1] = red[0].tix;
        o[2] = red[0].tiy;
        o[3] = red[0].tiz;
        o[4] = red[0].bad;
    }
}

// out: [0] max|div|, [1] u_tau, [2] dti * dt, [3] dti, [4] non-finite entries.
__global__ __launch_bounds__(kDiagBlock) void diag_final_kernel(Grid g,
                                                                const double* __restrict__ u,
                                                                const double* __restrict__ part,
                                                                int nblocks, double dt,
                                                                double* __restrict__ out) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    __shared__ double sb[kDiagBlock], st[kDiagBlock];
    __shared__ DiagPart red[kDiagBlock];
    DiagPart m{0.0, 0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
        const double* o = part + 5 * (long)b;
        m.div = nanmax(m.div, o[0]);
        m.tix = nanmax(m.tix, o[1]);
        m.tiy = nanmax(m.tiy, o[2]);
        m.tiz = nanmax(m.tiz, o[3]);
        m.bad += o[4];
    }
    // plane sums of u[:, :, 1] and u[:, :, nz] over (nx+1) x (ny+2) entries
    double s_bot = 0.0, s_top = 0.0;
    const long ncol = (long)(nx + 1) * (ny + 2);
    for (long c = threadIdx.x; c < ncol; c += blockDim.x) {
        s_bot += u[c * (nz + 2) + 1];
        s_top += u[c * (nz + 2) + nz];
    }
    red[threadIdx.x] = m;
    sb[threadIdx.x] = s_bot;
    st[threadIdx.x] = s_top;
    __syncthreads();
    for (int s = kDiagBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            DiagPart& x = red[threadIdx.x];
            const DiagPart& y = red[threadIdx.x + s];
            x.div = nanmax(x.div, y.div);
            x.tix = nanmax(x.tix, y.tix);
            x.tiy = nanmax(x.tiy, y.tiy);
            x.tiz = nanmax(x.tiz, y.tiz);
            x.bad += y.bad;
            sb[threadIdx.x] += sb[threadIdx.x + s];
            st[threadIdx.x] += st[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double dist = g.z_c[1];
        const double mean_bot = sb[0] / (double)ncol;
        const double u_tau_bot = sqrt(fabs(g.nu * mean_bot / dist));
        double u_tau = u_tau_bot;
        if (!g.top_neumann) {
            const double mean_top = st[0] / (double)ncol;
            const double u_tau_top = sqrt(fabs(g.nu * mean_top / dist));
            u_tau = 0.5 * (u_tau_bot + u_tau_top);
        }
        const double dti = nanmax(nanmax(red[0].tix, red[0].tiy), red[0].tiz);
        out[0] = red[0].div;
        out[1] = u_tau;
        out[2] = dti * dt;
        out[3] = dti;
        out[4] = red[0].bad;
    }
}

}  // namespace

extern "C" int slc_diagnostics(slc_plan* p, const double* u, const double* v, const double* w,
                               const double* pr, double dt, double* out, void* stream) {
    if (!p) return fail(SLC_EINVAL, "null plan");
    if (!u || !v || !w || !out) return fail(SLC_EINVAL, "null array");
    DeviceGuard dg(p->device);
    const long ncell = (long)p->nx * p->ny * p->nz;
    long nb = (ncell + kDiagBlock - 1) / kDiagBlock;
    if (nb > kDiagMaxBlocks) nb = kDiagMaxBlocks;
    if (nb < 1) nb = 1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(diag_partial_kernel, dim3((unsigned)nb), dim3(kDiagBlock), 0, s, p->g, u, v,
                       w, pr, p->d_dzf, p->d_dzc, p->diag_part);
    HIP_OK(hipGetLastError());
    hipLaunchKernelGGL(diag_final_kernel, dim3(1), dim3(kDiagBlock), 0, s, p->g, u, p->diag_part,
                       (int)nb, dt, out);
    HIP_OK(hipGetLastError());
    return SLC_OK;
}
