// slc_stats.hip -- on-device turbulence statistics (include/slc_stats.h).
//
// One snapshot of TurbulenceStats.accumulate_statistics
// (slchannel/turbstats.py:156-361) as five small launches plus, with
// spectra planes, three more:
//
//   pass 1   plane sums of the cell-centred u, v, w and of the u / v nodes
//            and w faces, per z level, as column-chunk partials;
//   reduce   partials -> plane means (fixed order);
//   pass 2   sums of squared node / face fluctuations about those means and
//            of the cell-centred u'w' product;
//   reduce   partials -> per-snapshot moments;
//   update   running sums += this snapshot's profiles (w'w' faces averaged
//            to centres, turbstats.py:229-231);
//   planes   fluctuation planes of the spectra pairs into a contiguous batch;
//   rocFFT   batched real-forward 2-D transforms (torch.fft.rfft2 layout);
//   fold     auto / cross spectra, both walls averaged, folded onto
//            kx > 0, ky > 0, added to the running sums
//            (turbstats.py:256-351).
//
// Layout of the passes: a workgroup is 64 z lanes x 4 column groups.  The z
// lanes make every load a coalesced 512-byte row segment; each thread walks
// an x-run of 16 cells for 4 y-rows, carrying the previous u face in a
// register so the cell-centred u costs one load.  Partials go to
// part[(chunk * NQ + q) * (nz + 1) + k] and are summed in chunk order by the
// reduce kernel, so the result does not depend on scheduling.
#include <rocfft/rocfft.h>

#include <cstring>
#include <vector>

#include "../../include/slc_stats.h"
#include "slc_plan.h"

using namespace slc;

namespace {

constexpr int kLanes = 64;  // z lanes per workgroup
constexpr int kSub = 4;     // column groups per workgroup
constexpr int kIB = 16;     // x cells per thread run
constexpr int kJB = 16;     // y rows per workgroup (kJB / kSub per thread)
constexpr int kNQ1 = 6;     // pass-1 sums: ucc, vcc, wcc, u node, v node, w face
constexpr int kNQ2 = 4;     // pass-2 sums: u'u' node, v'v' node, w'w' face, u'w' centre

struct Dims {
    int nx, ny, nz, nbx, nby;
};

// Plane sums over one workgroup's columns for z lane k (0..nz-1); lane
// nz-1 also sums the top w face (index nz), which it loads anyway.  mean:
// pass-1 means, (6, nz + 1).
template <int PASS>
__global__ __launch_bounds__(kLanes * kSub) void stats_pass_kernel(
    Dims d, const double* __restrict__ u, const double* __restrict__ v,
    const double* __restrict__ w, const double* __restrict__ mean, double* __restrict__ part) {
    constexpr int NQ = PASS == 1 ? kNQ1 : kNQ2;
    __shared__ double red[kSub][NQ][kLanes];
    __shared__ double red_top[kSub];
    const int lane = threadIdx.x % kLanes, sub = threadIdx.x / kLanes;
    const int k = blockIdx.x * kLanes + lane;
    const int nx = d.nx, ny = d.ny, nz = d.nz, nzf = nz + 1;
    const int i0 = blockIdx.y * kIB, i1 = min(nx, i0 + kIB);
    const int j0 = blockIdx.z * kJB, j1 = min(ny, j0 + kJB);
    const size_t su = (size_t)(ny + 2) * (nz + 2), sv = (size_t)(ny + 1) * (nz + 2),
                 sw = (size_t)(ny + 2) * (nz + 1);
    double s[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) s[q] = 0.0;
    double top = 0.0;  // lane nz-1: sums of the top face
    double mun = 0.0, mvn = 0.0, mwf = 0.0, mU = 0.0, mW = 0.0, mwt = 0.0;
    if (PASS == 2 && k < nz) {
        mwf = mean[5 * nzf + k];
        mU = mean[0 * nzf + k];
        mW = mean[2 * nzf + k];
        mun = mean[3 * nzf + k];
        mvn = mean[4 * nzf + k];
        if (k == nz - 1) mwt = mean[5 * nzf + nz];
    }
    for (int j = j0 + sub; j < j1; j += kSub) {
        // u[i0, j+1, k+1], v[i0+1, j, k+1], w[i0+1, j+1, k]
        const double* ur = u + ((size_t)i0 * (ny + 2) + (j + 1)) * (nz + 2) + (k + 1);
        const double* vr = v + ((size_t)(i0 + 1) * (ny + 1) + j) * (nz + 2) + (k + 1);
        const double* wr = w + ((size_t)(i0 + 1) * (ny + 2) + (j + 1)) * (nz + 1) + k;
        if (k < nz) {
            double up = ur[0];
            for (int i = i0; i < i1; ++i) {
                const size_t o = (size_t)(i - i0);
                const double un = ur[(o + 1) * su];
                const double v0 = vr[o * sv], v1 = vr[o * sv + (nz + 2)];
                const double w0 = wr[o * sw], w1 = wr[o * sw + 1];
                const double ucc = 0.5 * (up + un);
                const double wcc = 0.5 * (w0 + w1);
                if (PASS == 1) {
                    s[0] += ucc;
                    s[1] += 0.5 * (v0 + v1);
                    s[2] += wcc;
                    s[3] += un;
                    s[4] += v1;
                    s[5] += w0;
                } else {
                    const double du = un - mun, dv = v1 - mvn, dw = w0 - mwf;
                    s[0] += du * du;
                    s[1] += dv * dv;
                    s[2] += dw * dw;
                    s[3] += (ucc - mU) * (wcc - mW);
                }
                if (k == nz - 1) {
                    if (PASS == 1) {
                        top += w1;
                    } else {
                        const double dt = w1 - mwt;
                        top += dt * dt;
                    }
                }
                up = un;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) red[sub][q][lane] = s[q];
    if (k == nz - 1) red_top[sub] = top;
    __syncthreads();
    if (sub == 0 && k < nz) {
        const size_t chunk = (size_t)blockIdx.z * d.nbx + blockIdx.y;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double r = red[0][q][lane];
#pragma unroll
            for (int t = 1; t < kSub; ++t) r += red[t][q][lane];
            part[(chunk * NQ + q) * nzf + k] = r;
        }
        if (k == nz - 1) {
            double r = red_top[0];
            for (int t = 1; t < kSub; ++t) r += red_top[t];
            // the top face is quantity 5 (w face) in pass 1, 2 (w'w' face) in
            // pass 2; the other quantities have no entry there (zero)
            constexpr int QT = PASS == 1 ? 5 : 2;
#pragma unroll
            for (int q = 0; q < NQ; ++q) part[(chunk * NQ + q) * nzf + nz] = q == QT ? r : 0.0;
        }
    }
}

// Sums the partials of quantity q = blockIdx.y at lane k in chunk order
// (column group t of kRSub takes chunks t, t + kRSub, ...; groups combined in
// order) and divides by the plane size: out[q * (nz + 1) + k].
constexpr int kRSub = 16;

__global__ __launch_bounds__(kLanes * kRSub) void stats_reduce_kernel(
    const double* __restrict__ part, int nchunks, int NQ, int nz, double plane,
    double* __restrict__ out) {
    __shared__ double red[kRSub][kLanes];
    const int lane = threadIdx.x % kLanes, sub = threadIdx.x / kLanes;
    const int k = blockIdx.x * kLanes + lane, q = blockIdx.y, nzf = nz + 1;
    double r = 0.0;
    if (k <= nz)
        for (int c = sub; c < nchunks; c += kRSub) r += part[((size_t)c * NQ + q) * nzf + k];
    red[sub][lane] = r;
    __syncthreads();
    if (sub == 0 && k <= nz) {
        double t = red[0][lane];
        for (int g = 1; g < kRSub; ++g) t += red[g][lane];
        out[(size_t)q * nzf + k] = t / plane;
    }
}

// Running sums += this snapshot (turbstats.py:177, 195-200, 239-242).
__global__ __launch_bounds__(256) void stats_update_kernel(int nz, const double* __restrict__ mean,
                                                            const double* __restrict__ mom,
                                                            double* __restrict__ prof) {
    const int nzf = nz + 1;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nz; k += gridDim.x * blockDim.x) {
        const double U = mean[0 * nzf + k], V = mean[1 * nzf + k], W = mean[2 * nzf + k];
        prof[0 * nz + k] += U;
        prof[1 * nz + k] += mom[0 * nzf + k];
        prof[2 * nz + k] += mom[1 * nzf + k];
        prof[3 * nz + k] += 0.5 * (mom[2 * nzf + k] + mom[2 * nzf + k + 1]);
        prof[4 * nz + k] += mom[3 * nzf + k];
        prof[5 * nz + k] += V;
        prof[6 * nz + k] += W;
        prof[7 * nz + k] += U * U;
        prof[8 * nz + k] += V * V;
        prof[9 * nz + k] += W * W;
        prof[10 * nz + k] += U * W;
    }
}

// Fluctuation planes of the spectra pairs, batch index 6 * pair + c:
// c = 0, 1, 2 bottom u node, v node, cell-centred w; 3, 4, 5 the same at
// the top with w negated (turbstats.py:256-276).  Planes are (nx, ny), y
// fastest.
__global__ __launch_bounds__(256) void stats_planes_kernel(
    Dims d, const double* __restrict__ u, const double* __restrict__ v,
    const double* __restrict__ w, const double* __restrict__ mean, const int* __restrict__ kp,
    double* __restrict__ planes) {
    const int nx = d.nx, ny = d.ny, nz = d.nz, nzf = nz + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nx * ny) return;
    const int i = t / ny, j = t % ny;
    const int b = blockIdx.y, pair = b / 6, c = b % 6;
    const int k = kp[2 * pair + (c < 3 ? 0 : 1)];  // 1..nz
    const int kk = k - 1;
    double val;
    if (c % 3 == 0) {
        val = u[((size_t)(i + 1) * (ny + 2) + (j + 1)) * (nz + 2) + k] - mean[3 * nzf + kk];
    } else if (c % 3 == 1) {
        val = v[((size_t)(i + 1) * (ny + 1) + (j + 1)) * (nz + 2) + k] - mean[4 * nzf + kk];
    } else {
        const double* wc = w + ((size_t)(i + 1) * (ny + 2) + (j + 1)) * (nz + 1);
        val = 0.5 * (wc[kk] + wc[kk + 1]) - mean[2 * nzf + kk];
        if (c == 5) val = -val;
    }
    planes[(size_t)b * nx * ny + t] = val;
}

__device__ inline double auto_spec(double2 f, double norm) {
    return (f.x * f.x + f.y * f.y) / norm;
}
__device__ inline double cross_spec(double2 f, double2 g, double norm) {
    return (f.x * g.x + f.y * g.y) / norm;
}

// spec[pair][c][a][b] += fold(0.5 (E_bot + E_top)) with
// fold(E)[a, b] = E[a + 1, b + 1] + E[nx - 1 - a, b + 1] (turbstats.py:278-351).
__global__ __launch_bounds__(256) void stats_fold_kernel(Dims d, const double2* __restrict__ F,
                                                         double norm, double* __restrict__ spec) {
    const int nx = d.nx, ny = d.ny, nkx = nx / 2, nky = ny / 2, nyc = ny / 2 + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nkx * nky) return;
    const int a = t / nky, b = t % nky, pair = blockIdx.y;
    const size_t plane = (size_t)nx * nyc;
    const double2* Fp = F + (size_t)pair * 6 * plane;
    const size_t pos = (size_t)(a + 1) * nyc + (b + 1), neg = (size_t)(nx - 1 - a) * nyc + (b + 1);
    double* out = spec + (size_t)pair * 4 * nkx * nky + t;
    const size_t cs = (size_t)nkx * nky;
    double e[4][2];
    for (int side = 0; side < 2; ++side) {
        const size_t o = side == 0 ? pos : neg;
        const double2 ub = Fp[0 * plane + o], vb = Fp[1 * plane + o], wb = Fp[2 * plane + o];
        const double2 ut = Fp[3 * plane + o], vt = Fp[4 * plane + o], wt = Fp[5 * plane + o];
        e[0][side] = 0.5 * (auto_spec(ub, norm) + auto_spec(ut, norm));
        e[1][side] = 0.5 * (auto_spec(vb, norm) + auto_spec(vt, norm));
        e[2][side] = 0.5 * (auto_spec(wb, norm) + auto_spec(wt, norm));
        e[3][side] = 0.5 * (cross_spec(ub, wb, norm) + cross_spec(ut, wt, norm));
    }
    for (int c = 0; c < 4; ++c) out[c * cs] += e[c][0] + e[c][1];
}

}  // namespace

struct slc_stats {
    int device = 0;
    Dims d{};
    int nzc = 0, nzr = 0, nchunks = 0, nplanes = 0;
    double* part = nullptr;    // nchunks x 6 x (nz + 1)
    double* mean = nullptr;    // 6 x (nz + 1)
    double* mom = nullptr;     // 4 x (nz + 1)
    int* kp = nullptr;         // 2 x nplanes (k_bot, k_top)
    double* planes = nullptr;  // 6 nplanes x nx x ny
    double2* fhat = nullptr;   // 6 nplanes x nx x (ny/2 + 1)
    rocfft_plan fwd = nullptr;
    rocfft_execution_info info = nullptr;
    void* work = nullptr;
    std::vector<void*> allocs;
    int64_t bytes = 0;
};

namespace {

int stats_alloc(slc_stats* st, size_t bytes, void** out) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 8);
    if (e != hipSuccess)
        return fail(SLC_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    st->allocs.push_back(p);
    st->bytes += (int64_t)bytes;
    *out = p;
    return SLC_OK;
}

}  // namespace

extern "C" {

int slc_stats_destroy(slc_stats* st) {
    if (!st) return SLC_OK;
    DeviceGuard dg(st->device);
    (void)hipDeviceSynchronize();
    if (st->fwd) rocfft_plan_destroy(st->fwd);
    if (st->info) rocfft_execution_info_destroy(st->info);
    for (void* p : st->allocs) (void)hipFree(p);
    delete st;
    return SLC_OK;
}

int slc_stats_create(int device, int nx, int ny, int nz, int nplanes, const int* k_bot,
                     const int* k_top, slc_stats** out) {
    if (!out) return fail(SLC_EINVAL, "null argument");
    *out = nullptr;
    if (nx < 1 || ny < 1 || nz < 1)
        return fail(SLC_EINVAL, "grid too small (nx=%d ny=%d nz=%d)", nx, ny, nz);
    if ((double)(nx + 2) * (ny + 2) * (nz + 2) >= 2147483647.0)
        return fail(SLC_EINVAL, "grid too large ((nx+2)(ny+2)(nz+2) >= 2^31)");
    if (nplanes < 0) return fail(SLC_EINVAL, "nplanes must be >= 0");
    if (nplanes > 0) {
        if (!k_bot || !k_top) return fail(SLC_EINVAL, "null plane index arrays");
        if (nx < 2 || ny < 2) return fail(SLC_EINVAL, "spectra need nx, ny >= 2");
        for (int i = 0; i < nplanes; ++i)
            if (k_bot[i] < 1 || k_bot[i] > nz || k_top[i] < 1 || k_top[i] > nz)
                return fail(SLC_EINVAL, "plane %d: k_bot=%d k_top=%d outside 1..%d", i, k_bot[i],
                            k_top[i], nz);
    }
    DeviceGuard dg(device);
    slc_stats* st = new slc_stats();
    st->device = device;
    st->d = Dims{nx, ny, nz, (nx + kIB - 1) / kIB, (ny + kJB - 1) / kJB};
    st->nzc = (nz + kLanes - 1) / kLanes;        // pass workgroups along z
    st->nzr = (nz + 1 + kLanes - 1) / kLanes;    // reduce workgroups along z (faces)
    st->nchunks = st->d.nbx * st->d.nby;
    st->nplanes = nplanes;
    const size_t nzf = (size_t)nz + 1;
    void* p = nullptr;
    int rc;
    if ((rc = stats_alloc(st, sizeof(double) * st->nchunks * kNQ1 * nzf, &p))) goto bad;
    st->part = (double*)p;
    if ((rc = stats_alloc(st, sizeof(double) * kNQ1 * nzf, &p))) goto bad;
    st->mean = (double*)p;
    if ((rc = stats_alloc(st, sizeof(double) * kNQ2 * nzf, &p))) goto bad;
    st->mom = (double*)p;
    if (nplanes > 0) {
        std::vector<int> kp(2 * (size_t)nplanes);
        for (int i = 0; i < nplanes; ++i) {
            kp[2 * i] = k_bot[i];
            kp[2 * i + 1] = k_top[i];
        }
        if ((rc = stats_alloc(st, sizeof(int) * kp.size(), &p))) goto bad;
        st->kp = (int*)p;
        hipError_t e = hipMemcpy(st->kp, kp.data(), sizeof(int) * kp.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            rc = fail(SLC_EHIP, "plane table upload: %s", hipGetErrorString(e));
            goto bad;
        }
        const size_t nb = 6 * (size_t)nplanes;
        if ((rc = stats_alloc(st, sizeof(double) * nb * nx * ny, &p))) goto bad;
        st->planes = (double*)p;
        if ((rc = stats_alloc(st, sizeof(double2) * nb * nx * (ny / 2 + 1), &p))) goto bad;
        st->fhat = (double2*)p;
        rocfft_setup();
        // torch.fft.rfft2 over (x, y): y is the halved dimension (fastest first)
        size_t lengths[2] = {(size_t)ny, (size_t)nx};
        if (rocfft_plan_create(&st->fwd, rocfft_placement_notinplace,
                               rocfft_transform_type_real_forward, rocfft_precision_double, 2,
                               lengths, nb, nullptr) != rocfft_status_success) {
            rc = fail(SLC_EFFT, "rocfft_plan_create(spectra, batch %zu) failed", nb);
            goto bad;
        }
        if (rocfft_execution_info_create(&st->info) != rocfft_status_success) {
            rc = fail(SLC_EFFT, "rocfft_execution_info_create failed");
            goto bad;
        }
        size_t wb = 0;
        rocfft_plan_get_work_buffer_size(st->fwd, &wb);
        if (wb) {
            if ((rc = stats_alloc(st, wb, &p))) goto bad;
            st->work = p;
            rocfft_execution_info_set_work_buffer(st->info, st->work, wb);
        }
    }
    *out = st;
    return SLC_OK;
bad:
    slc_stats_destroy(st);
    return rc;
}

int slc_stats_workspace_bytes(const slc_stats* st, int64_t* bytes) {
    if (!st || !bytes) return fail(SLC_EINVAL, "null argument");
    *bytes = st->bytes;
    return SLC_OK;
}

int slc_stats_accumulate(slc_stats* st, const double* u, const double* v, const double* w,
                         double* prof, double* spec, void* stream) {
    if (!st) return fail(SLC_EINVAL, "null stats");
    if (!u || !v || !w || !prof) return fail(SLC_EINVAL, "null array");
    if (st->nplanes > 0 && !spec) return fail(SLC_EINVAL, "spec required with %d planes",
                                              st->nplanes);
    DeviceGuard dg(st->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const Dims d = st->d;
    const dim3 grid(st->nzc, d.nbx, d.nby), blk(kLanes * kSub);
    const double plane = (double)d.nx * (double)d.ny;
    hipLaunchKernelGGL(stats_pass_kernel<1>, grid, blk, 0, s, d, u, v, w, st->mean, st->part);
    HIP_OK(hipGetLastError());
    hipLaunchKernelGGL(stats_reduce_kernel, dim3(st->nzr, kNQ1), dim3(kLanes * kRSub), 0, s,
                       st->part, st->nchunks,
                       kNQ1, d.nz, plane, st->mean);
    hipLaunchKernelGGL(stats_pass_kernel<2>, grid, blk, 0, s, d, u, v, w, st->mean, st->part);
    hipLaunchKernelGGL(stats_reduce_kernel, dim3(st->nzr, kNQ2), dim3(kLanes * kRSub), 0, s,
                       st->part, st->nchunks,
                       kNQ2, d.nz, plane, st->mom);
    hipLaunchKernelGGL(stats_update_kernel, dim3((d.nz + 255) / 256), dim3(256), 0, s, d.nz,
                       st->mean, st->mom, prof);
    HIP_OK(hipGetLastError());
    if (st->nplanes == 0) return SLC_OK;
    const int npts = d.nx * d.ny;
    hipLaunchKernelGGL(stats_planes_kernel, dim3((npts + 255) / 256, 6 * st->nplanes), dim3(256),
                       0, s, d, u, v, w, st->mean, st->kp, st->planes);
    HIP_OK(hipGetLastError());
    FFT_OK(rocfft_execution_info_set_stream(st->info, s));
    void* in[1] = {st->planes};
    void* outp[1] = {st->fhat};
    FFT_OK(rocfft_execute(st->fwd, in, outp, st->info));
    const int nfold = (d.nx / 2) * (d.ny / 2);
    if (nfold > 0) {
        hipLaunchKernelGGL(stats_fold_kernel, dim3((nfold + 255) / 256, st->nplanes), dim3(256), 0,
                           s, d, st->fhat, plane * plane, spec);
        HIP_OK(hipGetLastError());
    }
    return SLC_OK;
}

}  // extern "C"
