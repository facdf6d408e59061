// slc_plan.h -- the plan object behind slc_plan* and the host helpers the
// library's C-ABI sources share (error text, HIP status checks, device guard,
// tracked device allocations).
#pragma once

#include <rocfft/rocfft.h>

#include <cstdint>
#include <vector>

#include "../../include/slc.h"
#include "slc_internal.h"

namespace slc {

// Sets the calling thread's slc_last_error() text; returns code.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

#define HIP_OK(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(SLC_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                               \
    } while (0)

#define FFT_OK(expr)                                                                   \
    do {                                                                               \
        rocfft_status s_ = (expr);                                                     \
        if (s_ != rocfft_status_success)                                               \
            return fail(SLC_EFFT, "%s failed: rocfft status %d", #expr, (int)s_);     \
    } while (0)

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) == hipSuccess && prev != dev) (void)hipSetDevice(dev);
        else prev = -1;
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

}  // namespace slc

struct slc_plan {
    using Grid = slc::Grid;
    using PoissonTabs = slc::PoissonTabs;
    using ArrivalTabs = slc::ArrivalTabs;
    int device = 0;
    hipStream_t init_stream = nullptr;  // non-blocking stream of the table uploads
    Grid g{};
    PoissonTabs pt{};
    int nx = 0, ny = 0, nz = 0, nky = 0;
    double total_volume = 0.0;
    std::vector<void*> allocs;
    int64_t bytes = 0;
    const double* d_dzc = nullptr;
    const double* d_dzf = nullptr;
    const double* d_dzf_inv = nullptr;  // 1 / dz_f
    const double* d_cv = nullptr;
    double* planes = nullptr;   // (nz, nx, ny) real
    double* dhat = nullptr;     // (nz, nx, nky) complex
    double* cpw = nullptr;      // (nz, nx, nky) Thomas c'
    double* ztab = nullptr;     // per component: 3 * (nz + 2) z-solve factor
    double zkey[3][2] = {};     // (dt, theta) the factor of each component was built for
    bool zvalid[3] = {false, false, false};
    double* partials = nullptr;
    int max_partials = 0;
    double* diag_part = nullptr;  // slc_diagnostics partials (5 per workgroup)
    double* us[3] = {nullptr, nullptr, nullptr};  // U* = 2V^n - V^{n-1} (ghost-shaped)
    double* nfoot = nullptr;                      // near feet of one component (3 x N)
    unsigned* defer = nullptr;  // slc_gather's per-tile flags for the wide second pass
    ArrivalTabs at[3] = {};
    // opt-in launch timing: hipEvents recorded around each labelled launch
    bool prof = false;
    int gather_shape = 0;  // -1: direct global loads; >= 0: LDS tile shape
    int bdf2_fused = 1;    // 1: one pass per component (MODE_BDF2); 0: far then near pass
    int storage = 64;      // 64: gathers read FP64 fields; 32: float copies (f32n / f32m)
    float* f32n[3] = {nullptr, nullptr, nullptr};  // float copy of V^n per component
    float* f32m[3] = {nullptr, nullptr, nullptr};  // float copy of V^{n-1} per component
    struct Rec {
        const char* label;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    rocfft_plan fwd = nullptr, inv = nullptr;
    rocfft_execution_info info = nullptr;
    void* fft_work = nullptr;
    size_t fft_work_bytes = 0;
};

namespace slc {

// hipMalloc tracked by the plan (freed by slc_plan_destroy).
int dev_alloc(slc_plan* p, size_t bytes, void** out);

// Workgroups of the diagnostics pass (slc_diagnostics); the plan holds
// 5 doubles of partials per workgroup.
constexpr int kDiagMaxBlocks = 2048;

// Plan construction; spectral = false skips the Poisson planes, spectra and
// rocFFT plans (members of an ensemble share per-size-class ones).
int plan_create(const slc_grid_desc* d, int device, bool spectral, slc_plan** out);

}  // namespace slc

namespace slc {

inline hipEvent_t pooled_event(slc_plan* p) {
    if (!p->pool.empty()) {
        hipEvent_t e = p->pool.back();
        p->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Records events around one launch when profiling is on (slc_profile_enable).
struct Timed {
    slc_plan* p;
    const char* label;
    hipStream_t s;
    hipEvent_t a = nullptr;
    Timed(slc_plan* p_, const char* l, void* st)
        : p(p_), label(l), s(static_cast<hipStream_t>(st)) {
        if (p->prof && (a = pooled_event(p))) (void)hipEventRecord(a, s);
    }
    ~Timed() {
        if (!a) return;
        hipEvent_t b = pooled_event(p);
        if (!b) return;
        (void)hipEventRecord(b, s);
        p->recs.push_back({label, a, b});
    }
};

}  // namespace slc
