// slc_api.hip -- plan object and the extern "C" entry points of libslc_hip.so.
//
// The plan owns only small tables and the spectral work buffers; fields are
// borrowed device pointers.  Host-side table construction follows the
// reference formulas operation for operation (cited per table) so the HIP
// path and the CPU oracle see the same numbers.
#include <rocfft/rocfft.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "slc_plan.h"

using namespace slc;

namespace slc {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int dev_alloc(slc_plan* p, size_t bytes, void** out) {
    void* ptr = nullptr;
    hipError_t e = hipMalloc(&ptr, bytes ? bytes : 8);
    if (e != hipSuccess)
        return fail(SLC_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    p->allocs.push_back(ptr);
    p->bytes += (int64_t)bytes;
    *out = ptr;
    return SLC_OK;
}

}  // namespace slc

namespace {

std::once_flag g_fft_once;

// Table upload on the plan's creation stream (non-blocking: it neither waits
// for nor delays work other callers have queued on the device).
int upload(slc_plan* p, const std::vector<double>& h, const double** out) {
    void* d = nullptr;
    int rc = dev_alloc(p, h.size() * sizeof(double), &d);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                          p->init_stream));
    HIP_OK(hipStreamSynchronize(p->init_stream));
    *out = static_cast<const double*>(d);
    return SLC_OK;
}

// semilag.py:102-114: inv_denom[b, m] = 1 / prod_l (z[b+m] - z[b+l] + [l == m]),
// product accumulated in increasing l.
std::vector<double> denom_table(const std::vector<double>& zn, int order) {
    int nb = (int)zn.size() - order + 1;
    std::vector<double> t((size_t)(nb > 0 ? nb : 0) * order);
    for (int b = 0; b < nb; ++b)
        for (int m = 0; m < order; ++m) {
            double prod = 0.0;
            for (int l = 0; l < order; ++l) {
                double d = (zn[b + m] - zn[b + l]) + (l == m ? 1.0 : 0.0);
                prod = (l == 0) ? d : prod * d;
            }
            t[(size_t)b * order + m] = 1.0 / prod;
        }
    return t;
}

// torch.fft.fftfreq / rfftfreq with d = delta / (2 pi), then the modified
// wavenumber (2/delta) sin(k delta / 2) squared (projection.py:47-60).
std::vector<double> modified_k2(int n, double delta, bool half) {
    const double two_pi = 2 * 3.141592653589793;
    double d = delta / two_pi;
    double val = 1.0 / ((double)n * d);
    int m = half ? n / 2 + 1 : n;
    std::vector<double> out(m);
    for (int i = 0; i < m; ++i) {
        int f = (half || i < (n + 1) / 2) ? i : i - n;
        double k = (double)f * val;
        double km = (2.0 / delta) * std::sin(k * delta / 2.0);
        out[i] = km * km;
    }
    return out;
}

int check_plan(slc_plan* p) {
    if (!p) return fail(SLC_EINVAL, "null plan");
    return SLC_OK;
}

inline hipStream_t S(void* s) { return static_cast<hipStream_t>(s); }

const char* kLabel[3][6] = {
    {"bdf2_far_u", "bdf2_near_u", "bdf1_u", "advect_u", "gather_u", "departure_u"},
    {"bdf2_far_v", "bdf2_near_v", "bdf1_v", "advect_v", "gather_v", "departure_v"},
    {"bdf2_far_w", "bdf2_near_w", "bdf1_w", "advect_w", "gather_w", "departure_w"},
};
const char* kZLabel[3] = {"zsolve_u", "zsolve_v", "zsolve_w"};
const char* kBdf2Label[3] = {"bdf2_u", "bdf2_v", "bdf2_w"};

int fft_exec(slc_plan* p, bool forward, void* stream) {
    FFT_OK(rocfft_execution_info_set_stream(p->info, stream));
    void* in[1] = {forward ? (void*)p->planes : (void*)p->dhat};
    void* out[1] = {forward ? (void*)p->dhat : (void*)p->planes};
    FFT_OK(rocfft_execute(forward ? p->fwd : p->inv, in, out, p->info));
    return SLC_OK;
}

// planes (nz, nx, ny) -> Poisson solution in planes (unnormalised)
int poisson_planes(slc_plan* p, void* stream) {
    int rc;
    {
        Timed t(p, "fft_forward", stream);
        rc = fft_exec(p, true, stream);
    }
    if (rc) return rc;
    {
        Timed t(p, "poisson_thomas", stream);
        HIP_OK(run_poisson_thomas(p->pt, p->dhat, p->cpw, S(stream)));
    }
    Timed t(p, "fft_inverse", stream);
    return fft_exec(p, false, stream);
}

}  // namespace

int slc::plan_create(const slc_grid_desc* d, int device, bool spectral, slc_plan** out) {
    if (!d || !out) return fail(SLC_EINVAL, "null argument");
    *out = nullptr;
    if (d->nx < 1 || d->ny < 2 || d->nz < 2)
        return fail(SLC_EINVAL, "grid too small (nx=%d ny=%d nz=%d)", d->nx, d->ny, d->nz);
    if (d->order != 4 && d->order != 6)
        return fail(SLC_EINVAL, "sl.interp_order must be 4 or 6, got %d", d->order);
    if (d->traj_order != 2 && d->traj_order != 4)
        return fail(SLC_EINVAL,
                    "sl.traj_interp_order must be 2 (trilinear) or 4 (tricubic), got %d",
                    d->traj_order);
    if (d->nz + 1 < d->traj_order)
        return fail(SLC_EINVAL, "nz=%d too small for trajectory order %d", d->nz,
                    d->traj_order);
    if (d->nz + 1 < d->order)
        return fail(SLC_EINVAL, "nz=%d too small for interpolation order %d", d->nz, d->order);
    if (d->stretching != 0 && d->stretching != 1)
        return fail(SLC_EINVAL, "stretching must be 0 (symmetric) or 1 (bottom)");
    if (d->n_traj_iters < 0) return fail(SLC_EINVAL, "n_traj_iters must be >= 0");
    if (!d->z_f || !d->z_c) return fail(SLC_EINVAL, "z_f / z_c host tables required");
    if ((double)(d->nx + 2) * (d->ny + 2) * (d->nz + 2) >= 2147483647.0)
        return fail(SLC_EINVAL, "grid too large for one plan ((nx+2)(ny+2)(nz+2) >= 2^31)");

    DeviceGuard dg(device);
    std::call_once(g_fft_once, [] { rocfft_setup(); });
    slc_plan* p = new slc_plan();
    p->device = device;
    {
        hipError_t e = hipStreamCreateWithFlags(&p->init_stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete p;
            return fail(SLC_EHIP, "plan stream: %s", hipGetErrorString(e));
        }
    }
    const int nx = d->nx, ny = d->ny, nz = d->nz, P = d->order;
    p->nx = nx;
    p->ny = ny;
    p->nz = nz;
    p->nky = ny / 2 + 1;
    const double dx = d->Lx / nx, dy = d->Ly / ny;

    std::vector<double> zf(d->z_f, d->z_f + nz + 1), zc(d->z_c, d->z_c + nz + 2);
    std::vector<double> dzf(nz), dzc(nz + 1), cv(nz);
    for (int k = 0; k < nz; ++k) dzf[k] = zf[k + 1] - zf[k];      // utils.py:58
    for (int k = 0; k <= nz; ++k) dzc[k] = zc[k + 1] - zc[k];     // utils.py:59
    for (int k = 0; k < nz; ++k) cv[k] = dx * dy * dzf[k];        // solver.py:317
    p->total_volume = d->Lx * d->Ly * d->Lz;                      // solver.py:321

    Grid& g = p->g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.order = P;
    g.n_iters = d->n_traj_iters;
    g.symmetric = d->stretching == 0 ? 1 : 0;
    g.top_neumann = d->top_wall == 1 ? 1 : 0;
    g.xy_lagged = 0;
    g.dx = dx;
    g.dy = dy;
    g.Lz = d->Lz;
    g.gamma = d->gamma;
    g.tanh_g = d->tanh_g;
    g.rLz_f = (float)(1.0 / d->Lz);
    g.rgamma_f = (float)(1.0 / d->gamma);
    g.tanh_g_f = (float)d->tanh_g;
    g.split_f = (float)(g.symmetric ? -d->gamma / (2.0 * nz) : d->gamma / (4.0 * nz));
    const double eps_z = 1e-12 * d->Lz;  // semilag.py:245-247
    g.z_lo = zf[0] + eps_z;
    g.z_hi = zf[nz] - eps_z;
    g.nu = d->nu;

    int rc = 0;
    const double* tmp = nullptr;
#define UP(vec, dst)                         \
    do {                                     \
        rc = upload(p, (vec), &tmp);         \
        if (rc) {                            \
            slc_plan_destroy(p);             \
            return rc;                       \
        }                                    \
        dst = tmp;                           \
    } while (0)
    UP(zc, g.z_c);
    UP(zf, g.z_f);
    UP(denom_table(zc, 2), g.dinv_c2);
    UP(denom_table(zf, 2), g.dinv_f2);
    UP(denom_table(zc, P), g.dinv_cP);
    UP(denom_table(zf, P), g.dinv_fP);
    g.traj_order = d->traj_order;
    if (g.traj_order == 2) {
        g.dinv_cT = g.dinv_c2;
        g.dinv_fT = g.dinv_f2;
    } else if (g.traj_order == P) {
        g.dinv_cT = g.dinv_cP;
        g.dinv_fT = g.dinv_fP;
    } else {
        UP(denom_table(zc, g.traj_order), g.dinv_cT);
        UP(denom_table(zf, g.traj_order), g.dinv_fT);
    }
    UP(dzc, p->d_dzc);
    UP(dzf, p->d_dzf);
    {
        // 1 / dz_f as the reference's tensor reciprocal (operators.py:669)
        std::vector<double> dzfi(dzf.size());
        for (size_t k = 0; k < dzf.size(); ++k) dzfi[k] = 1.0 / dzf[k];
        UP(dzfi, p->d_dzf_inv);
    }
    UP(cv, p->d_cv);

    // Poisson tables (projection.py:65-103) as 1-D factors
    std::vector<double> cl(nz), cr(nz), cs(nz);
    for (int k = 0; k < nz; ++k) {
        cl[k] = 1.0 / (dzc[k] * dzf[k]);
        cr[k] = 1.0 / (dzc[k + 1] * dzf[k]);
        cs[k] = cl[k] + cr[k];
    }
    PoissonTabs& pt = p->pt;
    pt.nx = nx;
    pt.ny = ny;
    pt.nz = nz;
    pt.nky = p->nky;
    UP(modified_k2(nx, dx, false), pt.kxsq);
    UP(modified_k2(ny, dy, true), pt.kysq);
    UP(cl, pt.cl);
    UP(cr, pt.cr);
    UP(cs, pt.cs);
#undef UP

    const long nplane = (long)nx * ny, nmodes = (long)nx * p->nky;
    void* ptr = nullptr;
    if (spectral) {
        if ((rc = dev_alloc(p, sizeof(double) * nplane * nz, &ptr))) { slc_plan_destroy(p); return rc; }
        p->planes = (double*)ptr;
        if ((rc = dev_alloc(p, 2 * sizeof(double) * nmodes * nz, &ptr))) { slc_plan_destroy(p); return rc; }
        p->dhat = (double*)ptr;
        if ((rc = dev_alloc(p, sizeof(double) * nmodes * nz, &ptr))) { slc_plan_destroy(p); return rc; }
        p->cpw = (double*)ptr;
    }
    if ((rc = dev_alloc(p, sizeof(double) * 9 * (nz + 2), &ptr))) { slc_plan_destroy(p); return rc; }
    p->ztab = (double*)ptr;
    {
        const size_t fsz[3] = {(size_t)(nx + 1) * (ny + 2) * (nz + 2),
                               (size_t)(nx + 2) * (ny + 1) * (nz + 2),
                               (size_t)(nx + 2) * (ny + 2) * (nz + 1)};
        for (int c = 0; c < 3; ++c) {
            if ((rc = dev_alloc(p, sizeof(double) * fsz[c], &ptr))) { slc_plan_destroy(p); return rc; }
            p->us[c] = (double*)ptr;
        }
        if ((rc = dev_alloc(p, sizeof(double) * 3 * (size_t)nx * ny * nz, &ptr))) {
            slc_plan_destroy(p);
            return rc;
        }
        p->nfoot = (double*)ptr;
        {
            long tiles = 0;
            for (int c = 0; c < 3; ++c) tiles = std::max(tiles, gather_tiles(g, c));
            if ((rc = dev_alloc(p, sizeof(unsigned) * (size_t)tiles, &ptr))) {
                slc_plan_destroy(p);
                return rc;
            }
            p->defer = (unsigned*)ptr;
        }
        for (int c = 0; c < 3; ++c) {
            const int nzi = c == 2 ? nz - 1 : nz;
            void *a = nullptr, *b = nullptr, *q = nullptr;
            if ((rc = dev_alloc(p, sizeof(double) * nx, &a)) ||
                (rc = dev_alloc(p, sizeof(double) * ny, &b)) ||
                (rc = dev_alloc(p, sizeof(int) * (nzi > 0 ? nzi : 1), &q))) {
                slc_plan_destroy(p);
                return rc;
            }
            p->at[c] = ArrivalTabs{(const double*)a, (const double*)b, (const int*)q};
            hipError_t e = run_arrival_tabs(g, c, (double*)a, (double*)b, (int*)q,
                                            p->init_stream);
            if (e == hipSuccess) e = hipStreamSynchronize(p->init_stream);
            if (e != hipSuccess) {
                slc_plan_destroy(p);
                return fail(SLC_EHIP, "arrival tables: %s", hipGetErrorString(e));
            }
        }
    }
    p->max_partials = bulk_partials_needed(g);
    if ((rc = dev_alloc(p, sizeof(double) * p->max_partials, &ptr))) { slc_plan_destroy(p); return rc; }
    p->partials = (double*)ptr;
    if ((rc = dev_alloc(p, sizeof(double) * 5 * kDiagMaxBlocks, &ptr))) { slc_plan_destroy(p); return rc; }
    p->diag_part = (double*)ptr;

    if (!spectral) {
        *out = p;
        return SLC_OK;
    }
    // rocFFT: nz batched 2-D transforms of contiguous (nx, ny) planes; the
    // halved (Hermitian) dimension is y, as torch.fft.rfft2 over dims (0, 1).
    size_t lengths[2] = {(size_t)ny, (size_t)nx};  // fastest first
    rocfft_status fs;
    fs = rocfft_plan_create(&p->fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                            rocfft_precision_double, 2, lengths, (size_t)nz, nullptr);
    if (fs != rocfft_status_success) {
        slc_plan_destroy(p);
        return fail(SLC_EFFT, "rocfft_plan_create(forward) status %d", (int)fs);
    }
    fs = rocfft_plan_create(&p->inv, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                            rocfft_precision_double, 2, lengths, (size_t)nz, nullptr);
    if (fs != rocfft_status_success) {
        slc_plan_destroy(p);
        return fail(SLC_EFFT, "rocfft_plan_create(inverse) status %d", (int)fs);
    }
    size_t w1 = 0, w2 = 0;
    rocfft_plan_get_work_buffer_size(p->fwd, &w1);
    rocfft_plan_get_work_buffer_size(p->inv, &w2);
    p->fft_work_bytes = w1 > w2 ? w1 : w2;
    rocfft_execution_info_create(&p->info);
    if (p->fft_work_bytes) {
        if ((rc = dev_alloc(p, p->fft_work_bytes, &ptr))) { slc_plan_destroy(p); return rc; }
        p->fft_work = ptr;
        rocfft_execution_info_set_work_buffer(p->info, p->fft_work, p->fft_work_bytes);
    }
    *out = p;
    return SLC_OK;
}

// ---------------------------------------------------------------------------
// slc_bulk_force_noalloc: the bulk-flux shift with its partial sums in a
// file-scope device array instead of the plan's workspace.  Each device holds
// its own copy of `scratch`; its address on a device is looked up once, with
// that device current, and kept in g_scratch_dev[].  Host code uses only that
// address: the array's name in host code is not a device address.
// ---------------------------------------------------------------------------

namespace {
constexpr int kNoAllocPartials = 2048;  // the most bulk_partials_needed() returns
constexpr int kNoAllocDevices = 64;
}  // namespace

static __device__ __align__(256) double scratch[kNoAllocPartials];

namespace {
std::mutex g_scratch_mu;
double* g_scratch_dev[kNoAllocDevices] = {};

// The device address of `scratch` on `device` (current here), looked up on
// the first call for that device.
int scratch_on(int device, double** out) {
    if (device < 0 || device >= kNoAllocDevices)
        return fail(SLC_EINVAL, "device %d outside 0..%d", device, kNoAllocDevices - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(SLC_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                        hipGetErrorString(e));
        g_scratch_dev[device] = static_cast<double*>(ptr);
    }
    *out = g_scratch_dev[device];
    return SLC_OK;
}
}  // namespace

extern "C" {

int slc_version(void) { return 100; }

const char* slc_last_error(void) { return g_err.c_str(); }

int slc_plan_destroy(slc_plan* p) {
    if (!p) return SLC_OK;
    DeviceGuard dg(p->device);
    if (p->fwd) rocfft_plan_destroy(p->fwd);
    if (p->inv) rocfft_plan_destroy(p->inv);
    if (p->info) rocfft_execution_info_destroy(p->info);
    if (p->init_stream) (void)hipStreamDestroy(p->init_stream);
    for (auto& r : p->recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (hipEvent_t e : p->pool) (void)hipEventDestroy(e);
    for (void* a : p->allocs) (void)hipFree(a);
    delete p;
    return SLC_OK;
}

int slc_plan_create(const slc_grid_desc* d, int device, slc_plan** out) {
    return plan_create(d, device, true, out);
}

int slc_plan_workspace_bytes(const slc_plan* p, int64_t* out) {
    if (!p || !out) return fail(SLC_EINVAL, "null argument");
    *out = p->bytes;
    return SLC_OK;
}

int slc_apply_bc(slc_plan* p, double* u, double* v, double* w, void* stream) {
    if (int rc = check_plan(p)) return rc;
    DeviceGuard dg(p->device);
    Timed t(p, "ghost_fill", stream);
    HIP_OK(run_bc(p->g, u, v, w, S(stream)));
    return SLC_OK;
}

int slc_departure(slc_plan* p, int comp, const double* um, const double* vm, const double* wm,
                  int layout, double dt, double* xd, double* yd, double* zd, int64_t* nclamp,
                  void* stream) {
    if (int rc = check_plan(p)) return rc;
    if (comp < 0 || comp > 2) return fail(SLC_EINVAL, "comp must be 0, 1 or 2");
    if (layout != 0 && layout != 1) return fail(SLC_EINVAL, "layout must be 0 or 1");
    if (!um || !vm || !wm || !xd || !yd || !zd) return fail(SLC_EINVAL, "null array");
    DeviceGuard dg(p->device);
    TrajVel V{{um, vm, wm}, {nullptr, nullptr, nullptr}};
    Timed t(p, kLabel[comp][5], stream);
    HIP_OK(run_departure(p->g, comp, V, layout, dt, xd, yd, zd,
                         reinterpret_cast<unsigned long long*>(nclamp), S(stream)));
    return SLC_OK;
}

int slc_gather(slc_plan* p, int comp, const double* field, int layout, const double* xd,
               const double* yd, const double* zd, double* out, int64_t* nclamp, void* stream) {
    if (int rc = check_plan(p)) return rc;
    if (comp < 0 || comp > 2) return fail(SLC_EINVAL, "comp must be 0, 1 or 2");
    if (layout != 0 && layout != 1) return fail(SLC_EINVAL, "layout must be 0 or 1");
    if (!field || !xd || !yd || !zd || !out) return fail(SLC_EINVAL, "null array");
    DeviceGuard dg(p->device);
    long fsx = 0, fsy = 0;
    if (layout == 1) {
        long NZ = comp == 2 ? p->nz + 1 : p->nz + 2;
        fsy = NZ;
        fsx = (long)p->ny * NZ;
    }
    Timed t(p, kLabel[comp][4], stream);
    if (layout == 0 && p->gather_shape == 0) {
        TrajVel V{{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
        HIP_OK(run_tile_step(p->g, comp, 4, p->at[comp], V, field, nullptr, nullptr, nullptr,
                             0.0, out,
                             const_cast<double*>(xd), const_cast<double*>(yd),
                             const_cast<double*>(zd),
                             reinterpret_cast<unsigned long long*>(nclamp), S(stream),
                             p->defer));
    } else if (layout == 0 && p->gather_shape > 0)
        HIP_OK(run_gather_tiled(p->g, comp, p->gather_shape, field, xd, yd, zd, out,
                                reinterpret_cast<unsigned long long*>(nclamp), S(stream)));
    else
        HIP_OK(run_gather(p->g, comp, field, fsx, fsy, 0, 0, xd, yd, zd, out,
                          reinterpret_cast<unsigned long long*>(nclamp), S(stream)));
    return SLC_OK;
}

int slc_stencil(slc_plan* p, int comp, const double* xd, const double* yd, const double* zd,
                int32_t* ix, int32_t* iy, int32_t* k0, int64_t* nclamp, void* stream) {
    if (int rc = check_plan(p)) return rc;
    if (comp < 0 || comp > 2) return fail(SLC_EINVAL, "comp must be 0, 1 or 2");
    if (!xd || !yd || !zd || !ix || !iy || !k0) return fail(SLC_EINVAL, "null array");
    DeviceGuard dg(p->device);
    HIP_OK(run_stencil(p->g, comp, xd, yd, zd, ix, iy, k0,
                       reinterpret_cast<unsigned long long*>(nclamp), S(stream)));
    return SLC_OK;
}

int slc_locate_z(slc_plan* p, int comp, const double* z, int64_t n, int32_t* m0_exact,
                 int32_t* m0_fast, void* stream) {
    if (int rc = check_plan(p)) return rc;
    if (comp < 0 || comp > 2) return fail(SLC_EINVAL, "comp must be 0, 1 or 2");
    if (!z || !m0_exact || !m0_fast || n < 0) return fail(SLC_EINVAL, "bad array");
    DeviceGuard dg(p->device);
    HIP_OK(run_locate(p->g, comp, z, (long)n, m0_exact, m0_fast, S(stream)));
    return SLC_OK;
}

// The LDS-tiled step kernels trace with trilinear trajectories; tricubic ones
// (traj_interp_order 4) run the direct-load kernels.
static bool tiled_steps(const slc_plan* p) {
    return p->gather_shape >= 0 && p->g.traj_order == 2;
}

int slc_advect(slc_plan* p, const double* u, const double* v, const double* w, const double* um,
               const double* vm, const double* wm, double dt, double* ustar, double* vstar,
               double* wstar, int64_t* nclamp, void* stream) {
    if (int rc = check_plan(p)) return rc;
    if (!u || !v || !w || !um || !vm || !wm || !ustar || !vstar || !wstar)
        return fail(SLC_EINVAL, "null array");
    DeviceGuard dg(p->device);
    TrajVel V{{um, vm, wm}, {nullptr, nullptr, nullptr}};
    const double* F[3] = {u, v, w};
    double* O[3] = {ustar, vstar, wstar};
    auto* ncl = reinterpret_cast<unsigned long long*>(nclamp);
    const bool f32 = p->storage == 32 && tiled_steps(p);
    if (f32) {
        Timed t(p, "to_f32", stream);
        HIP_OK(run_to_f32(p->g, p->f32n, F, S(stream)));
    }
    for (int c = 0; c < 3; ++c) {
        Timed t(p, kLabel[c][3], stream);
        if (tiled_steps(p))
            HIP_OK(run_tile_step(p->g, c, 0, p->at[c], V, F[c], nullptr,
                                 f32 ? p->f32n[c] : nullptr, nullptr, dt, O[c], nullptr,
                                 nullptr, nullptr, ncl, S(stream)));
        else
            HIP_OK(run_advect(p->g, c, V, F[c], dt, O[c], ncl, S(stream)));
    }
    return SLC_OK;
}

int slc_bdf1_predict(slc_plan* p, const double* u, const double* v, const double* w, double dt,
                     double* ustar, double* vstar, double* wstar, int64_t* nclamp,
                     void* stream) {
    if (int rc = check_plan(p)) return rc;
    if (!u || !v || !w || !ustar || !vstar || !wstar) return fail(SLC_EINVAL, "null array");
    DeviceGuard dg(p->device);
    TrajVel V{{u, v, w}, {nullptr, nullptr, nullptr}};
    const double* F[3] = {u, v, w};
    double* O[3] = {ustar, vstar, wstar};
    auto* ncl = reinterpret_cast<unsigned long long*>(nclamp);
    const bool f32 = p->storage == 32 && tiled_steps(p);
    if (f32) {
        Timed t(p, "to_f32", stream);
        HIP_OK(run_to_f32(p->g, p->f32n, F, S(stream)));
    }
    for (int c = 0; c < 3; ++c) {
        Timed t(p, kLabel[c][2], stream);
        if (tiled_steps(p))
            HIP_OK(run_tile_step(p->g, c, 1, p->at[c], V, F[c], nullptr,
                                 f32 ? p->f32n[c] : nullptr, nullptr, dt, O[c], nullptr,
                                 nullptr, nullptr, ncl, S(stream)));
        else
            HIP_OK(run_bdf1(p->g, c, V, F[c], dt, O[c], ncl, S(stream)));
    }
    return SLC_OK;
}

int slc_bdf2_predict(slc_plan* p, const double* u, const double* v, const double* w,
                     const double* unm1, const double* vnm1, const double* wnm1, double dt,
                     double* ustar, double* vstar, double* wstar, int64_t* nclamp,
                     void* stream) {
    if (int rc = check_plan(p)) return rc;
    if (!u || !v || !w || !unm1 || !vnm1 || !wnm1 || !ustar || !vstar || !wstar)
        return fail(SLC_EINVAL, "null array");
    DeviceGuard dg(p->device);
    const double* Fn[3] = {u, v, w};
    const double* Fm[3] = {unm1, vnm1, wnm1};
    double* O[3] = {ustar, vstar, wstar};
    auto* ncl = reinterpret_cast<unsigned long long*>(nclamp);
    if (tiled_steps(p)) {
        const bool f32 = p->storage == 32;
        {
            Timed t(p, "ustar", stream);
            HIP_OK(run_ustar(p->g, p->us, Fn, Fm, f32 ? p->f32n : nullptr,
                             f32 ? p->f32m : nullptr, S(stream)));
        }
        TrajVel V{{p->us[0], p->us[1], p->us[2]}, {nullptr, nullptr, nullptr}};
        const long N = (long)p->nx * p->ny * p->nz;
        if (p->bdf2_fused) {
            for (int c = 0; c < 3; ++c) {
                Timed t(p, kBdf2Label[c], stream);
                HIP_OK(run_tile_step(p->g, c, 5, p->at[c], V, Fn[c], Fm[c],
                                     f32 ? p->f32n[c] : nullptr, f32 ? p->f32m[c] : nullptr, dt,
                                     O[c], nullptr, nullptr, nullptr, ncl, S(stream)));
            }
            return SLC_OK;
        }
        for (int c = 0; c < 3; ++c)
            for (int pass = 0; pass < 2; ++pass) {
                Timed t(p, kLabel[c][pass], stream);
                HIP_OK(run_tile_step(p->g, c, 2 + pass, p->at[c], V, Fn[c], Fm[c],
                                     f32 ? p->f32n[c] : nullptr, f32 ? p->f32m[c] : nullptr, dt,
                                     O[c], p->nfoot, p->nfoot + N, p->nfoot + 2 * N, ncl,
                                     S(stream)));
            }
        return SLC_OK;
    }
    TrajVel V{{u, v, w}, {unm1, vnm1, wnm1}};
    for (int pass = 0; pass < 2; ++pass)
        for (int c = 0; c < 3; ++c) {
            Timed t(p, kLabel[c][pass], stream);
            HIP_OK(run_bdf2(p->g, c, pass, V, Fn[c], Fm[c], dt, O[c], ncl, S(stream)));
        }
    return SLC_OK;
}

int slc_zsolve(slc_plan* p, int comp, double* field, double dt, double theta, void* stream) {
    if (int rc = check_plan(p)) return rc;
    if (comp < 0 || comp > 2) return fail(SLC_EINVAL, "comp must be 0, 1 or 2");
    if (!field) return fail(SLC_EINVAL, "null array");
    if (!(theta > 0.0 && theta <= 1.0)) return fail(SLC_EINVAL, "theta must be in (0, 1]");
    if (comp == 2 && p->nz < 3) return fail(SLC_EINVAL, "w z-solve needs nz >= 3");
    DeviceGuard dg(p->device);
    // the factorisation depends only on (dt, theta): rebuilt when they change
    bool refactor = !(p->zvalid[comp] && p->zkey[comp][0] == dt && p->zkey[comp][1] == theta);
    p->zvalid[comp] = true;
    p->zkey[comp][0] = dt;
    p->zkey[comp][1] = theta;
    Timed t(p, kZLabel[comp], stream);
    HIP_OK(run_zsolve(p->g, comp, field, dt, theta, p->d_dzc, p->d_dzf,
                      p->ztab + (size_t)comp * 3 * (p->nz + 2), refactor, S(stream)));
    return SLC_OK;
}

int slc_divergence(slc_plan* p, const double* u, const double* v, const double* w, double* div,
                   void* stream) {
    if (int rc = check_plan(p)) return rc;
    if (!u || !v || !w || !div) return fail(SLC_EINVAL, "null array");
    DeviceGuard dg(p->device);
    HIP_OK(run_div_field(p->g, u, v, w, p->d_dzf, div, S(stream)));
    return SLC_OK;
}

int slc_poisson(slc_plan* p, const double* div, double rhs_scale, double* pr, void* stream) {
    if (int rc = check_plan(p)) return rc;
    if (!div || !pr) return fail(SLC_EINVAL, "null array");
    DeviceGuard dg(p->device);
    HIP_OK(run_field_to_planes(p->g, div, rhs_scale, p->planes, S(stream)));
    if (int rc = poisson_planes(p, stream)) return rc;
    const double inv_n = 1.0 / ((double)p->nx * (double)p->ny);
    HIP_OK(run_project_planes(p->g, nullptr, nullptr, nullptr, p->planes, inv_n, 0.0, p->d_dzc,
                              pr, S(stream)));
    HIP_OK(run_bc_p(p->g, pr, S(stream)));
    return SLC_OK;
}

int slc_project(slc_plan* p, double* u, double* v, double* w, const double* pr, double dt,
                void* stream) {
    if (int rc = check_plan(p)) return rc;
    if (!u || !v || !w || !pr) return fail(SLC_EINVAL, "null array");
    DeviceGuard dg(p->device);
    HIP_OK(run_project_p(p->g, u, v, w, pr, dt, p->d_dzc, S(stream)));
    return SLC_OK;
}

int slc_pressure_combine(slc_plan* p, double* out, double a, const double* x, double b,
                         const double* y, void* stream) {
    if (int rc = check_plan(p)) return rc;
    if (!out || !x || !y) return fail(SLC_EINVAL, "null array");
    DeviceGuard dg(p->device);
    const long n = (long)(p->nx + 2) * (p->ny + 2) * (p->nz + 2);
    HIP_OK(run_axpby(out, x, y, a, b, n, S(stream)));
    return SLC_OK;
}

int slc_project_step(slc_plan* p, double* u, double* v, double* w, double rhs_scale,
                     double dt_proj, double* pr, void* stream) {
    if (int rc = check_plan(p)) return rc;
    if (!u || !v || !w) return fail(SLC_EINVAL, "null array");
    DeviceGuard dg(p->device);
    {
        Timed t(p, "divergence_planes", stream);
        HIP_OK(run_div_planes(p->g, u, v, w, p->d_dzf, rhs_scale, p->planes, S(stream)));
    }
    if (int rc = poisson_planes(p, stream)) return rc;
    const double inv_n = 1.0 / ((double)p->nx * (double)p->ny);
    {
        Timed t(p, "project_planes", stream);
        HIP_OK(run_project_planes(p->g, u, v, w, p->planes, inv_n, dt_proj, p->d_dzc, pr,
                                  S(stream)));
    }
    Timed t(p, "ghost_fill_p_uvw", stream);
    if (pr) HIP_OK(run_bc_p(p->g, pr, S(stream)));
    HIP_OK(run_bc(p->g, u, v, w, S(stream)));
    return SLC_OK;
}

int slc_bulk_velocity(slc_plan* p, const double* u, double* out, void* stream) {
    if (int rc = check_plan(p)) return rc;
    if (!u || !out) return fail(SLC_EINVAL, "null array");
    DeviceGuard dg(p->device);
    HIP_OK(run_bulk_velocity(p->g, u, p->d_cv, p->total_volume, p->partials, p->max_partials, out,
                             S(stream)));
    return SLC_OK;
}

int slc_bulk_force(slc_plan* p, double* u, double* v, double* w, double U_bulk, double dt,
                   double* out, void* stream) {
    if (int rc = check_plan(p)) return rc;
    if (!u || !out) return fail(SLC_EINVAL, "null array");
    if (!(dt != 0.0)) return fail(SLC_EINVAL, "dt must be non-zero");
    DeviceGuard dg(p->device);
    Timed t(p, "bulk_force", stream);
    HIP_OK(run_bulk(p->g, u, v, w, p->d_cv, p->total_volume, U_bulk, dt, 1, p->partials,
                    p->max_partials, out, S(stream)));
    return SLC_OK;
}

int slc_bulk_force_noalloc(slc_plan* p, double* u, double* v, double* w, double U_bulk,
                           double dt, double* out, void* stream) {
    if (int rc = check_plan(p)) return rc;
    if (!u || !out) return fail(SLC_EINVAL, "null array");
    if (!(dt != 0.0)) return fail(SLC_EINVAL, "dt must be non-zero");
    const int np = bulk_partials_needed(p->g);
    if (np > kNoAllocPartials)
        return fail(SLC_EINVAL, "%d bulk partial sums exceed the static scratch of %d", np,
                    kNoAllocPartials);
    DeviceGuard dg(p->device);
    double* partials = nullptr;
    if (int rc = scratch_on(p->device, &partials)) return rc;
    // every partial is written before it is read: nothing to clear
    HIP_OK(run_bulk(p->g, u, v, w, p->d_cv, p->total_volume, U_bulk, dt, 1, partials,
                    kNoAllocPartials, out, S(stream)));
    return SLC_OK;
}

int slc_set_viscosity(slc_plan* p, double nu) {
    if (int rc = check_plan(p)) return rc;
    if (!(nu > 0.0) || !std::isfinite(nu)) return fail(SLC_EINVAL, "nu must be positive, got %g", nu);
    p->g.nu = nu;
    for (int c = 0; c < 3; ++c) p->zvalid[c] = false;  // the z-solve factors hold nu
    return SLC_OK;
}

int slc_set_option(slc_plan* p, int option, int value) {
    if (int rc = check_plan(p)) return rc;
    switch (option) {
        case SLC_OPT_GATHER_TILE:
            if (value < -1 || value > 3) return fail(SLC_EINVAL, "gather tile shape %d", value);
            p->gather_shape = value;
            return SLC_OK;
        case SLC_OPT_BDF2_XY_RHS:
            if (value != 0 && value != 1) return fail(SLC_EINVAL, "bdf2 xy rhs must be 0 or 1");
            p->g.xy_lagged = value;
            return SLC_OK;
        case SLC_OPT_BDF2_FUSED:
            if (value != 0 && value != 1) return fail(SLC_EINVAL, "bdf2 fused must be 0 or 1");
            p->bdf2_fused = value;
            return SLC_OK;
        case SLC_OPT_FIELD_STORAGE: {
            if (value != 32 && value != 64)
                return fail(SLC_EINVAL, "field storage must be 32 or 64 bits, got %d", value);
            if (value == 32 && !p->f32n[0]) {
                DeviceGuard dg(p->device);
                const long n[3] = {(long)(p->nx + 1) * (p->ny + 2) * (p->nz + 2),
                                   (long)(p->nx + 2) * (p->ny + 1) * (p->nz + 2),
                                   (long)(p->nx + 2) * (p->ny + 2) * (p->nz + 1)};
                for (int c = 0; c < 3; ++c) {
                    void* a = nullptr;
                    void* b = nullptr;
                    if (int rc = dev_alloc(p, (size_t)n[c] * sizeof(float), &a)) return rc;
                    if (int rc = dev_alloc(p, (size_t)n[c] * sizeof(float), &b)) return rc;
                    p->f32n[c] = static_cast<float*>(a);
                    p->f32m[c] = static_cast<float*>(b);
                }
            }
            p->storage = value;
            return SLC_OK;
        }
        default:
            return fail(SLC_EINVAL, "unknown option %d", option);
    }
}

int slc_profile_enable(slc_plan* p, int on) {
    if (int rc = check_plan(p)) return rc;
    p->prof = on != 0;
    return SLC_OK;
}

int slc_profile_report(slc_plan* p, char* buf, int64_t len) {
    if (int rc = check_plan(p)) return rc;
    if (!buf || len <= 0) return fail(SLC_EINVAL, "null/empty buffer");
    DeviceGuard dg(p->device);
    struct Acc {
        std::string label;
        double total = 0.0;
        long n = 0;
    };
    std::vector<Acc> acc;
    for (auto& r : p->recs) {
        HIP_OK(hipEventSynchronize(r.b));
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, r.a, r.b));
        size_t i = 0;
        for (; i < acc.size(); ++i)
            if (acc[i].label == r.label) break;
        if (i == acc.size()) acc.push_back({r.label, 0.0, 0});
        acc[i].total += ms;
        acc[i].n += 1;
        p->pool.push_back(r.a);
        p->pool.push_back(r.b);
    }
    p->recs.clear();
    std::string out;
    char line[160];
    for (auto& a : acc) {
        snprintf(line, sizeof line, "%s %ld %.6f\n", a.label.c_str(), a.n, a.total);
        out += line;
    }
    if ((int64_t)out.size() + 1 > len) return fail(SLC_EINVAL, "report buffer too small");
    memcpy(buf, out.c_str(), out.size() + 1);
    return SLC_OK;
}

}  // extern "C"
