// slc_imex.hip -- the Eulerian IMEX explicit terms (slc_imex_rhs, slc_ab2_update).
//
// compute_momentum_rhs_fused_imex (operators.py:639-835; the Triton kernel of
// eulerian_triton.py:31-193 computes the same): divergence-form advection on
// the MAC grid and the xy Laplacian, rhs = diffusion_xy - advection on each
// component's interior and 0 on every other entry of its ghost-shaped array.
// The three components in one launch; each output element uses the
// reference's operation order (no contraction), so it is the same double the
// eager path computes.  The neighbours come from L1/L2: the HBM traffic is the
// three fields once plus the three outputs.
//
// slc_ab2_update: u += dt (1.5 r - 0.5 r_prev) (forward Euler, u += dt r,
// without a previous rhs) over the whole arrays (solver.py:1074-1081).
#include "slc_plan.h"

using namespace slc;

namespace {

struct Rhs3 {
    double *u, *v, *w;
};

struct ImexK {
    double dx_inv, dy_inv, nu_dx2, nu_dy2;
};

// One thread per entry (i, j, k) of the common (nx+2, ny+2, nz+2) index
// space (z fastest, blocks kept on one XCD in order); it writes the u, v and w
// entries (i, j, k) where the component's array has one.
__global__ __launch_bounds__(256) void imex_rhs_kernel(Grid g, const double* __restrict__ u,
                                                       const double* __restrict__ v,
                                                       const double* __restrict__ w,
                                                       const double* __restrict__ dzfi,
                                                       const double* __restrict__ dzc, ImexK c,
                                                       Rhs3 out) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const unsigned t = (unsigned)xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= (unsigned)((nx + 2) * (ny + 2) * (nz + 2))) return;
    const int k = (int)(t % (unsigned)(nz + 2));
    const unsigned r = t / (unsigned)(nz + 2);
    const int j = (int)(r % (unsigned)(ny + 2)), i = (int)(r / (unsigned)(ny + 2));
    // strides of the three arrays (z fastest)
    const int uz = nz + 2, uy = (ny + 2) * uz;  // u: (nx+1, ny+2, nz+2)
    const int vz = nz + 2, vy = (ny + 1) * vz;  // v: (nx+2, ny+1, nz+2)
    const int wz = nz + 1, wy = (ny + 2) * wz;  // w: (nx+2, ny+2, nz+1)
    {
        // ---- u (operators.py:679-728); U(i+1) at i = nx is the periodic u[1]
        if (i <= nx) {
            const int o = i * uy + j * uz + k;
            double res = 0.0;
            if (i >= 1 && j >= 1 && j <= ny && k >= 1 && k <= nz) {
                const int ip = (i == nx) ? o - (nx - 1) * uy : o + uy;  // U(i+1)
                const int im = o - uy;
                const int vo = i * vy + j * vz + k;  // v[i, j, k]
                const int wo = i * wy + j * wz + k;  // w[i, j, k]
                const double a = 0.5 * (u[o] + u[ip]);
                const double b = 0.5 * (u[im] + u[o]);
                const double duudx = (a * a - b * b) * c.dx_inv;
                double vi = 0.5 * (v[vo] + v[vo + vy]);
                double ui = 0.5 * (u[o] + u[o + uz]);
                const double vu_top = vi * ui;
                vi = 0.5 * (v[vo - vz] + v[vo + vy - vz]);
                ui = 0.5 * (u[o - uz] + u[o]);
                const double dvudy = (vu_top - vi * ui) * c.dy_inv;
                double wi = 0.5 * (w[wo] + w[wo + wy]);
                ui = 0.5 * (u[o] + u[o + 1]);
                const double wu_top = wi * ui;
                wi = 0.5 * (w[wo - 1] + w[wo + wy - 1]);
                ui = 0.5 * (u[o - 1] + u[o]);
                const double dwudz = (wu_top - wi * ui) * dzfi[k - 1];
                const double adv = duudx + dvudy + dwudz;
                const double d2x = (u[ip] - 2.0 * u[o] + u[im]) * c.nu_dx2;
                const double d2y = (u[o + uz] - 2.0 * u[o] + u[o - uz]) * c.nu_dy2;
                res = (d2x + d2y) - adv;
            }
            out.u[o] = res;
        }
        // ---- v (operators.py:734-781); V(j+1) at j = ny is the periodic v[:, 1]
        if (j <= ny) {
            const int o = i * vy + j * vz + k;
            double res = 0.0;
            if (i >= 1 && i <= nx && j >= 1 && k >= 1 && k <= nz) {
                const int jp = (j == ny) ? o - (ny - 1) * vz : o + vz;  // V(j+1)
                const int uo = i * uy + j * uz + k;  // u[i, j, k]
                const int wo = i * wy + j * wz + k;  // w[i, j, k]
                double ui = 0.5 * (u[uo] + u[uo + uz]);
                double vi = 0.5 * (v[o] + v[o + vy]);
                const double uv_right = ui * vi;
                ui = 0.5 * (u[uo - uy] + u[uo - uy + uz]);
                vi = 0.5 * (v[o - vy] + v[o]);
                const double duvdx = (uv_right - ui * vi) * c.dx_inv;
                const double a = 0.5 * (v[o] + v[jp]);
                const double b = 0.5 * (v[o - vz] + v[o]);
                const double dvvdy = (a * a - b * b) * c.dy_inv;
                double wi = 0.5 * (w[wo] + w[wo + wz]);
                vi = 0.5 * (v[o] + v[o + 1]);
                const double wv_top = wi * vi;
                wi = 0.5 * (w[wo - 1] + w[wo + wz - 1]);
                vi = 0.5 * (v[o - 1] + v[o]);
                const double dwvdz = (wv_top - wi * vi) * dzfi[k - 1];
                const double adv = duvdx + dvvdy + dwvdz;
                const double d2x = (v[o + vy] - 2.0 * v[o] + v[o - vy]) * c.nu_dx2;
                const double d2y = (v[jp] - 2.0 * v[o] + v[o - vz]) * c.nu_dy2;
                res = (d2x + d2y) - adv;
            }
            out.v[o] = res;
        }
        // ---- w (operators.py:787-833), interior faces k = 1..nz-1
        if (k <= nz) {
            const int o = i * wy + j * wz + k;
            double res = 0.0;
            if (i >= 1 && i <= nx && j >= 1 && j <= ny && k >= 1 && k <= nz - 1) {
                const int uo = i * uy + j * uz + k;  // u[i, j, k]
                const int vo = i * vy + j * vz + k;  // v[i, j, k]
                double ui = 0.5 * (u[uo] + u[uo + 1]);
                double wi = 0.5 * (w[o] + w[o + wy]);
                const double uw_right = ui * wi;
                ui = 0.5 * (u[uo - uy] + u[uo - uy + 1]);
                wi = 0.5 * (w[o - wy] + w[o]);
                const double duwdx = (uw_right - ui * wi) * c.dx_inv;
                double vi = 0.5 * (v[vo] + v[vo + 1]);
                wi = 0.5 * (w[o] + w[o + wz]);
                const double vw_top = vi * wi;
                vi = 0.5 * (v[vo - vz] + v[vo - vz + 1]);
                wi = 0.5 * (w[o - wz] + w[o]);
                const double dvwdy = (vw_top - vi * wi) * c.dy_inv;
                const double a = 0.5 * (w[o] + w[o + 1]);
                const double b = 0.5 * (w[o - 1] + w[o]);
                const double dwwdz = (a * a - b * b) / dzc[k];
                const double adv = duwdx + dvwdy + dwwdz;
                const double d2x = (w[o + wy] - 2.0 * w[o] + w[o - wy]) * c.nu_dx2;
                const double d2y = (w[o + wz] - 2.0 * w[o] + w[o - wz]) * c.nu_dy2;
                res = (d2x + d2y) - adv;
            }
            out.w[o] = res;
        }
    }
}

__global__ __launch_bounds__(256) void ab2_update_kernel(double* __restrict__ u,
                                                         double* __restrict__ v,
                                                         double* __restrict__ w, Rhs3 r,
                                                         Rhs3 rp, double dt, unsigned nu_,
                                                         unsigned nv, unsigned nw) {
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    double* f[3] = {u, v, w};
    const double* a[3] = {r.u, r.v, r.w};
    const double* b[3] = {rp.u, rp.v, rp.w};
    const unsigned n[3] = {nu_, nv, nw};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (t >= n[c]) continue;
        if (b[c]) f[c][t] = f[c][t] + dt * (1.5 * a[c][t] - 0.5 * b[c][t]);
        else f[c][t] = f[c][t] + dt * a[c][t];
    }
}

}  // namespace

extern "C" int slc_imex_rhs(slc_plan* p, const double* u, const double* v, const double* w,
                            double* ru, double* rv, double* rw, void* stream) {
    if (!p) return fail(SLC_EINVAL, "null plan");
    if (!u || !v || !w || !ru || !rv || !rw) return fail(SLC_EINVAL, "null array");
    DeviceGuard dg(p->device);
    const Grid& g = p->g;
    // operators.py:662-672: reciprocals and nu-scaled squares as the reference forms them
    ImexK c;
    c.dx_inv = 1.0 / g.dx;
    c.dy_inv = 1.0 / g.dy;
    c.nu_dx2 = g.nu * (1.0 / (g.dx * g.dx));
    c.nu_dy2 = g.nu * (1.0 / (g.dy * g.dy));
    Timed tm(p, "imex_rhs", stream);
    const long ncom = (long)(g.nx + 2) * (g.ny + 2) * (g.nz + 2);
    hipLaunchKernelGGL(imex_rhs_kernel, dim3((unsigned)((ncom + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), g, u, v, w, p->d_dzf_inv, p->d_dzc, c,
                       Rhs3{ru, rv, rw});
    HIP_OK(hipGetLastError());
    return SLC_OK;
}

extern "C" int slc_ab2_update(slc_plan* p, double* u, double* v, double* w, const double* ru,
                              const double* rv, const double* rw, const double* ru_prev,
                              const double* rv_prev, const double* rw_prev, double dt,
                              void* stream) {
    if (!p) return fail(SLC_EINVAL, "null plan");
    if (!u || !v || !w || !ru || !rv || !rw) return fail(SLC_EINVAL, "null array");
    const bool prev = ru_prev || rv_prev || rw_prev;
    if (prev && !(ru_prev && rv_prev && rw_prev))
        return fail(SLC_EINVAL, "previous rhs: all three components or none");
    DeviceGuard dg(p->device);
    const Grid& g = p->g;
    const unsigned n[3] = {(unsigned)((g.nx + 1) * (g.ny + 2) * (g.nz + 2)),
                           (unsigned)((g.nx + 2) * (g.ny + 1) * (g.nz + 2)),
                           (unsigned)((g.nx + 2) * (g.ny + 2) * (g.nz + 1))};
    const unsigned nmax = n[0] > n[1] ? (n[0] > n[2] ? n[0] : n[2]) : (n[1] > n[2] ? n[1] : n[2]);
    Timed tm(p, "ab2_update", stream);
    hipLaunchKernelGGL(ab2_update_kernel, dim3((nmax + 255) / 256), dim3(256), 0,
                       static_cast<hipStream_t>(stream), u, v, w,
                       Rhs3{(double*)ru, (double*)rv, (double*)rw},
                       Rhs3{(double*)ru_prev, (double*)rv_prev, (double*)rw_prev}, dt, n[0],
                       n[1], n[2]);
    HIP_OK(hipGetLastError());
    return SLC_OK;
}
