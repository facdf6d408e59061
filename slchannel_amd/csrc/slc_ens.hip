// slc_ens.hip -- the batched ensemble (include/slc_ens.h): member arenas,
// size classes with shared spectral buffers and batched rocFFT plans, the
// device member table, one batched launch per substep, optional hipGraph
// replay, and the tile-shape autotuner.
#include <rocfft/rocfft.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/slc_ens.h"
#include "slc_ens.h"
#include "slc_ops_dev.h"
#include "slc_plan.h"

using namespace slc;

namespace {

// one member's host record
// everything of a grid description but the viscosity: members with equal keys
// can share one arena (tables, buffers) one after the other
struct GeoKey {
    int nx, ny, nz, stretching, order, traj_order, n_iters, top;
    double Lx, Ly, Lz, gamma, tanh_g;
    bool operator==(const GeoKey& o) const {
        return nx == o.nx && ny == o.ny && nz == o.nz && stretching == o.stretching &&
               order == o.order && traj_order == o.traj_order && n_iters == o.n_iters &&
               top == o.top && Lx == o.Lx && Ly == o.Ly && Lz == o.Lz && gamma == o.gamma &&
               tanh_g == o.tanh_g;
    }
};

GeoKey geo_key(const slc_grid_desc* d) {
    return GeoKey{d->nx,         d->ny, d->nz, d->stretching, d->order, d->traj_order,
                  d->n_traj_iters, d->top_wall, d->Lx, d->Ly, d->Lz, d->gamma, d->tanh_g};
}

struct Member {
    int id = 0;
    GeoKey key{};
    slc_plan* plan = nullptr;  // grid / Poisson / arrival tables, U*, feet, partials
    int nx = 0, ny = 0, nz = 0;
    double dt = 0.0, U_bulk = 1.0;
    std::vector<void*> allocs;
    int64_t bytes = 0;
    double* fld[3][3] = {};
    double* p = nullptr;
    float* f32n[3] = {};
    float* f32m[3] = {};
    double* ztab = nullptr;  // [mode][comp] factor tables
    double* scal = nullptr;
    double* stats = nullptr;
    unsigned long long* ncl = nullptr;
    int* state = nullptr;
    int host_state = 0;  // mirror of *state (advanced with every step)
    int cls = -1, slot = -1;
};

// members of one grid size share plane-major spectral buffers (one slot per
// member) and one batched rocFFT plan per direction over all the slots
struct SizeClass {
    int nx, ny, nz;
    int count = 0;     // members in the table
    int cap = 0;       // slots allocated
    double* planes = nullptr;
    double* dhat = nullptr;
    double* cpw = nullptr;
    std::map<int, rocfft_plan> fwd, inv;  // by batch (slots); only the current capacity's
};

constexpr int kFieldSets = 3;

long field_elems(int c, int nx, int ny, int nz) {
    if (c == 0) return (long)(nx + 1) * (ny + 2) * (nz + 2);
    if (c == 1) return (long)(nx + 2) * (ny + 1) * (nz + 2);
    return (long)(nx + 2) * (ny + 2) * (nz + 1);
}

long p_elems(int nx, int ny, int nz) { return (long)(nx + 2) * (ny + 2) * (nz + 2); }

long spectral_bytes(int nx, int ny, int nz) {
    const long nky = ny / 2 + 1;
    return 8L * nx * ny * nz + 16L * nx * nky * nz + 8L * nx * nky * nz;
}

}  // namespace

struct slc_ens {
    int device = 0;
    int order = 6;
    int storage = 64;
    int tile_shape = 0;
    int grid_per_cu = 0;  // 0: one workgroup per work unit; k: k persistent workgroups per CU
    bool use_graph = false;
    int bdf2_fused = 1;  // 1: one fused BDF2 tile pass per component; 0: far, then near pass
    int ncu = 256;
    hipStream_t es = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    int next_id = 1;
    std::vector<Member*> members;
    std::vector<SizeClass> classes;
    bool dirty = true;
    // device table
    EnsMember* d_mem = nullptr;
    size_t d_mem_cap = 0;
    unsigned* d_start = nullptr;
    size_t d_start_cap = 0;
    int* d_umap[3] = {nullptr, nullptr, nullptr};
    size_t d_umap_cap[3] = {0, 0, 0};
    std::vector<unsigned> h_start;
    unsigned totals[EK_COUNT] = {};
    size_t zs_lds[3] = {0, 0, 0};
    bool zs_short[3] = {false, false, false};  // every member: <= 2 unknowns per lane
    // rocFFT
    rocfft_execution_info info = nullptr;
    void* fft_work = nullptr;
    size_t fft_work_bytes = 0;
    // graph
    hipGraphExec_t gexec = nullptr;
    int64_t class_bytes = 0;
    // arenas of removed members, reused by later members of the same geometry
    std::vector<Member*> pool;
    size_t pool_max = 8;
    // pinned, double-buffered staging of the device table (async uploads
    // ordered on es behind the steps still reading the previous table)
    struct Staging {
        char* host = nullptr;
        size_t cap = 0;
        hipEvent_t done = nullptr;
    } stage[2];
    int stage_i = 0;
};

namespace {

int ens_check(const slc_ens* e) {
    if (!e) return fail(SLC_EINVAL, "null ensemble");
    return SLC_OK;
}

Member* find(slc_ens* e, int id) {
    for (Member* m : e->members)
        if (m->id == id) return m;
    return nullptr;
}

int m_alloc(Member* m, size_t bytes, void** out) {
    void* p = nullptr;
    hipError_t err = hipMalloc(&p, bytes ? bytes : 8);
    if (err != hipSuccess)
        return fail(SLC_ENOMEM, "hipMalloc(%zu) for an ensemble member failed: %s", bytes,
                    hipGetErrorString(err));
    m->allocs.push_back(p);
    m->bytes += (int64_t)bytes;
    *out = p;
    return SLC_OK;
}

void free_member(Member* m) {
    if (!m) return;
    for (void* p : m->allocs) (void)hipFree(p);
    if (m->plan) slc_plan_destroy(m->plan);
    delete m;
}

void drop_graph(slc_ens* e) {
    if (e->gexec) {
        (void)hipGraphExecDestroy(e->gexec);
        e->gexec = nullptr;
    }
}

void free_class(SizeClass& c) {
    for (auto& kv : c.fwd) rocfft_plan_destroy(kv.second);
    for (auto& kv : c.inv) rocfft_plan_destroy(kv.second);
    c.fwd.clear();
    c.inv.clear();
    if (c.planes) (void)hipFree(c.planes);
    if (c.dhat) (void)hipFree(c.dhat);
    if (c.cpw) (void)hipFree(c.cpw);
    c.planes = c.dhat = c.cpw = nullptr;
    c.cap = 0;
}

// The class's transforms: one forward and one inverse plan over all its
// slots (cap members x nz planes), so a step executes one plan per class and
// direction whatever the member count; slots beyond the count are
// transformed too and never read.  Plans of an earlier capacity are dropped.
int class_plans(slc_ens* e, SizeClass& c) {
    const int q = c.cap;
    if (c.fwd.count(q)) return SLC_OK;
    for (auto& kv : c.fwd) rocfft_plan_destroy(kv.second);
    for (auto& kv : c.inv) rocfft_plan_destroy(kv.second);
    c.fwd.clear();
    c.inv.clear();
    size_t lengths[2] = {(size_t)c.ny, (size_t)c.nx};
    rocfft_plan f = nullptr, b = nullptr;
    rocfft_status s = rocfft_plan_create(&f, rocfft_placement_notinplace,
                                         rocfft_transform_type_real_forward,
                                         rocfft_precision_double, 2, lengths,
                                         (size_t)q * c.nz, nullptr);
    if (s != rocfft_status_success)
        return fail(SLC_EFFT, "rocfft_plan_create(forward, batch %d) status %d", q * c.nz,
                    (int)s);
    s = rocfft_plan_create(&b, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                           rocfft_precision_double, 2, lengths, (size_t)q * c.nz, nullptr);
    if (s != rocfft_status_success) {
        rocfft_plan_destroy(f);
        return fail(SLC_EFFT, "rocfft_plan_create(inverse, batch %d) status %d", q * c.nz,
                    (int)s);
    }
    c.fwd[q] = f;
    c.inv[q] = b;
    size_t w1 = 0, w2 = 0;
    rocfft_plan_get_work_buffer_size(f, &w1);
    rocfft_plan_get_work_buffer_size(b, &w2);
    const size_t w = std::max(w1, w2);
    if (w > e->fft_work_bytes) {
        if (e->fft_work) (void)hipFree(e->fft_work);
        e->fft_work = nullptr;
        e->fft_work_bytes = 0;
        HIP_OK(hipMalloc(&e->fft_work, w));
        e->fft_work_bytes = w;
        FFT_OK(rocfft_execution_info_set_work_buffer(e->info, e->fft_work, w));
    }
    return SLC_OK;
}

// device buffer of at least n elements; grows geometrically (a reallocation
// frees the old buffer, which waits for the device)
template <typename T>
int grow(T** p, size_t* cap, size_t n) {
    if (n <= *cap && *p) return SLC_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    size_t want = std::max<size_t>(std::max<size_t>(n, 16), 2 * *cap);
    *cap = 0;
    HIP_OK(hipMalloc(p, want * sizeof(T)));
    *cap = want;
    return SLC_OK;
}

unsigned tiles_of(const Member* m, int comp, int shape) {
    int ti, tj, tk;
    ens_tile_shape(shape, ti, tj, tk);
    const int nzi = comp == 2 ? m->nz - 1 : m->nz;
    return (unsigned)(((m->nx + ti - 1) / ti) * ((m->ny + tj - 1) / tj) * ((nzi + tk - 1) / tk));
}

// rebuild classes, slots, the device member table, work prefixes and the
// tile-to-member maps (outside any capture)
int rebuild(slc_ens* e) {
    if (e->gexec) {
        HIP_OK(hipStreamSynchronize(e->es));  // the captured graph may still run
        drop_graph(e);
    }
    const int n = (int)e->members.size();
    // size classes: keep existing ones (their plans), reset counts
    for (auto& c : e->classes) c.count = 0;
    for (Member* m : e->members) {
        int ci = -1;
        for (size_t k = 0; k < e->classes.size(); ++k)
            if (e->classes[k].nx == m->nx && e->classes[k].ny == m->ny &&
                e->classes[k].nz == m->nz)
                ci = (int)k;
        if (ci < 0) {
            SizeClass c;
            c.nx = m->nx;
            c.ny = m->ny;
            c.nz = m->nz;
            e->classes.push_back(c);
            ci = (int)e->classes.size() - 1;
        }
        m->cls = ci;
        m->slot = e->classes[ci].count++;
    }
    e->class_bytes = 0;
    for (auto& c : e->classes) {
        if (c.count == 0) {
            free_class(c);
            continue;
        }
        if (c.count > c.cap) {
            // hipFree waits for the device, so in-flight steps finish first
            if (c.planes) (void)hipFree(c.planes);
            if (c.dhat) (void)hipFree(c.dhat);
            if (c.cpw) (void)hipFree(c.cpw);
            c.planes = c.dhat = c.cpw = nullptr;
            const int want = std::max(c.count, c.cap + (c.cap + 1) / 2);
            c.cap = 0;
            const long nky = c.ny / 2 + 1;
            const size_t per = (size_t)c.nx * c.ny * c.nz;
            const size_t perk = (size_t)c.nx * nky * c.nz;
            HIP_OK(hipMalloc(&c.planes, per * want * sizeof(double)));
            HIP_OK(hipMalloc(&c.dhat, 2 * perk * want * sizeof(double)));
            HIP_OK(hipMalloc(&c.cpw, perk * want * sizeof(double)));
            c.cap = want;
        }
        e->class_bytes += (int64_t)c.cap * spectral_bytes(c.nx, c.ny, c.nz);
        if (int rc = class_plans(e, c)) return rc;
    }
    // member table
    std::vector<EnsMember> tab(n);
    e->h_start.assign((size_t)EK_COUNT * (n + 1), 0u);
    std::vector<int> umap[3];
    for (int k = 0; k < 3; ++k) {
        e->zs_lds[k] = 0;
        e->zs_short[k] = true;
    }
    for (int i = 0; i < n; ++i) {
        Member* m = e->members[i];
        slc_plan* pl = m->plan;
        SizeClass& c = e->classes[m->cls];
        EnsMember& M = tab[i];
        std::memset(&M, 0, sizeof M);
        M.g = pl->g;
        for (int k = 0; k < 3; ++k) M.at[k] = pl->at[k];
        for (int s = 0; s < kFieldSets; ++s)
            for (int k = 0; k < 3; ++k) M.fld[s][k] = m->fld[s][k];
        M.p = m->p;
        for (int k = 0; k < 3; ++k) {
            M.us[k] = pl->us[k];
            M.f32n[k] = m->f32n[k];
            M.f32m[k] = m->f32m[k];
        }
        M.nfoot = pl->nfoot;
        M.dzc = pl->d_dzc;
        M.dzf = pl->d_dzf;
        M.cv = pl->d_cv;
        M.ztab = m->ztab;
        M.pt = pl->pt;
        const long nky = m->ny / 2 + 1;
        M.planes = c.planes + (size_t)m->slot * m->nx * m->ny * m->nz;
        M.dhat = c.dhat + 2 * (size_t)m->slot * m->nx * nky * m->nz;
        M.cpw = c.cpw + (size_t)m->slot * m->nx * nky * m->nz;
        M.partials = pl->partials;
        M.scal = m->scal;
        M.stats = m->stats;
        M.ncl = m->ncl;
        M.state = m->state;
        M.dt = m->dt;
        M.U_bulk = m->U_bulk;
        M.total_volume = pl->total_volume;
        M.inv_n = 1.0 / ((double)m->nx * (double)m->ny);
        M.npart = pl->max_partials;
        for (int k = 0; k < 3; ++k) {
            M.zch[k] = zsolve_scan_ch(k, m->nz);
            M.zcols[k] = M.zch[k] > 0 ? kZScanCols * zscan_k(M.zch[k]) : zsolve_cols(k, m->nz);
            e->zs_short[k] = e->zs_short[k] && (M.zch[k] == 1 || M.zch[k] == 2);
        }
        for (int k = 0; k < 4; ++k) M.nghost[k] = ghost_count(k, m->nx, m->ny, m->nz);
        // work units
        unsigned units[EK_COUNT];
        units[EK_BC3] = ublocks(M.nghost[0]) + ublocks(M.nghost[1]) + ublocks(M.nghost[2]);
        units[EK_BCP] = ublocks(M.nghost[3]);
        units[EK_BCU] = ublocks(M.nghost[0]);
        long nmax = 0;
        for (int k = 0; k < 3; ++k) nmax = std::max(nmax, field_elems(k, m->nx, m->ny, m->nz));
        units[EK_ELEM] = ublocks(nmax);
        for (int k = 0; k < 3; ++k) {
            units[EK_TILE0 + k] = tiles_of(m, k, e->tile_shape);
            const long ncols = (long)m->nx * m->ny;
            units[EK_ZS0 + k] = (unsigned)((ncols + M.zcols[k] - 1) / M.zcols[k]);
            if (M.zch[k] == 0)
                e->zs_lds[k] = std::max(e->zs_lds[k], (size_t)M.zcols[k] *
                                                          zsolve_ld(k, m->nz) * sizeof(double));
        }
        units[EK_PLANE] = (unsigned)plane_tiles(m->nx, m->ny, m->nz);
        units[EK_THOMAS] = spike_ok(m->nz)
                               ? (unsigned)(((long)m->nx * nky + spike_modes(m->nz) - 1) /
                                            spike_modes(m->nz))
                               : ublocks((long)m->nx * nky);
        units[EK_BULK] = (unsigned)M.npart;
        units[EK_ONE] = 1;
        for (int k = 0; k < EK_COUNT; ++k)
            e->h_start[(size_t)k * (n + 1) + i + 1] = e->h_start[(size_t)k * (n + 1) + i] + units[k];
        for (int k = 0; k < 3; ++k) umap[k].insert(umap[k].end(), units[EK_TILE0 + k], i);
    }
    for (int k = 0; k < EK_COUNT; ++k) e->totals[k] = e->h_start[(size_t)k * (n + 1) + n];
    if (n > 0) {
        if (int rc = grow(&e->d_mem, &e->d_mem_cap, (size_t)n)) return rc;
        if (int rc = grow(&e->d_start, &e->d_start_cap, e->h_start.size())) return rc;
        for (int k = 0; k < 3; ++k)
            if (int rc = grow(&e->d_umap[k], &e->d_umap_cap[k], umap[k].size())) return rc;
        // serialise into the pinned staging buffer not used by the last upload
        const size_t sz_mem = n * sizeof(EnsMember);
        const size_t sz_start = e->h_start.size() * sizeof(unsigned);
        size_t total = sz_mem + sz_start;
        for (int k = 0; k < 3; ++k) total += umap[k].size() * sizeof(int);
        auto& st = e->stage[e->stage_i];
        e->stage_i ^= 1;
        if (st.done) HIP_OK(hipEventSynchronize(st.done));
        if (total > st.cap) {
            const size_t want = std::max(total, 2 * st.cap) + 4096;
            if (st.host) (void)hipHostFree(st.host);
            st.host = nullptr;
            st.cap = 0;
            HIP_OK(hipHostMalloc((void**)&st.host, want, hipHostMallocDefault));
            st.cap = want;
        }
        if (!st.done) HIP_OK(hipEventCreateWithFlags(&st.done, hipEventDisableTiming));
        size_t off = 0;
        std::memcpy(st.host + off, tab.data(), sz_mem);
        HIP_OK(hipMemcpyAsync(e->d_mem, st.host + off, sz_mem, hipMemcpyHostToDevice, e->es));
        off += sz_mem;
        std::memcpy(st.host + off, e->h_start.data(), sz_start);
        HIP_OK(hipMemcpyAsync(e->d_start, st.host + off, sz_start, hipMemcpyHostToDevice, e->es));
        off += sz_start;
        for (int k = 0; k < 3; ++k) {
            const size_t sz = umap[k].size() * sizeof(int);
            std::memcpy(st.host + off, umap[k].data(), sz);
            HIP_OK(hipMemcpyAsync(e->d_umap[k], st.host + off, sz, hipMemcpyHostToDevice, e->es));
            off += sz;
        }
        HIP_OK(hipEventRecord(st.done, e->es));
    }
    e->dirty = false;
    return SLC_OK;
}

EnsLaunch launch(const slc_ens* e, int kind, int comp = -1) {
    const int n = (int)e->members.size();
    EnsLaunch L;
    L.mem = e->d_mem;
    L.start = e->d_start + (size_t)kind * (n + 1);
    L.nmem = n;
    L.total = e->totals[kind];
    L.umap = comp >= 0 ? e->d_umap[comp] : nullptr;
    return L;
}

unsigned pgrid(const slc_ens* e, int kind) {
    const unsigned t = e->totals[kind] ? e->totals[kind] : 1u;
    if (e->grid_per_cu == 0) return t;
    const unsigned cap = (unsigned)(e->ncu * e->grid_per_cu);
    return t < cap ? t : cap;
}

int fft_all(slc_ens* e, bool forward) {
    FFT_OK(rocfft_execution_info_set_stream(e->info, e->es));
    for (SizeClass& c : e->classes) {
        if (c.count == 0) continue;
        void* in[1] = {forward ? (void*)c.planes : (void*)c.dhat};
        void* out[1] = {forward ? (void*)c.dhat : (void*)c.planes};
        FFT_OK(rocfft_execute(forward ? c.fwd[c.cap] : c.inv[c.cap], in, out, e->info));
    }
    return SLC_OK;
}

// SL passes of the three components: one fused pass each (BDF2 members: both
// feet and gathers; bootstrap members: BDF1), or far, then near / bootstrap
int sl_passes(slc_ens* e) {
    const int f32 = e->storage == 32 ? 1 : 0;
    if (e->bdf2_fused) {
        for (int c = 0; c < 3; ++c) {
            EnsLaunch L = launch(e, EK_TILE0 + c, c);
            HIP_OK(run_ens_tile(L, c, 2, e->order, f32, e->tile_shape, L.total, e->es));
        }
        return SLC_OK;
    }
    for (int c = 0; c < 3; ++c)
        for (int pass = 0; pass < 2; ++pass) {
            EnsLaunch L = launch(e, EK_TILE0 + c, c);
            HIP_OK(run_ens_tile(L, c, pass, e->order, f32, e->tile_shape, L.total, e->es));
        }
    return SLC_OK;
}

// the whole step, enqueued on e->es
int enqueue_step(slc_ens* e) {
    hipStream_t s = e->es;
    HIP_OK(run_ens_bc3(launch(e, EK_BC3), 0, pgrid(e, EK_BC3), s));
    HIP_OK(run_ens_ustar(launch(e, EK_ELEM), pgrid(e, EK_ELEM), s));
    if (int rc = sl_passes(e)) return rc;
    HIP_OK(run_ens_bc3(launch(e, EK_BC3), 1, pgrid(e, EK_BC3), s));
    for (int c = 0; c < 3; ++c)
        HIP_OK(run_ens_zsolve(launch(e, EK_ZS0 + c), c, e->zs_lds[c], pgrid(e, EK_ZS0 + c),
                              e->zs_short[c], s));
    HIP_OK(run_ens_bc3(launch(e, EK_BC3), 1, pgrid(e, EK_BC3), s));
    HIP_OK(run_ens_div_planes(launch(e, EK_PLANE), pgrid(e, EK_PLANE), s));
    if (int rc = fft_all(e, true)) return rc;
    HIP_OK(run_ens_thomas(launch(e, EK_THOMAS), pgrid(e, EK_THOMAS), s));
    if (int rc = fft_all(e, false)) return rc;
    HIP_OK(run_ens_project(launch(e, EK_PLANE), pgrid(e, EK_PLANE), s));
    HIP_OK(run_ens_bcp(launch(e, EK_BCP), pgrid(e, EK_BCP), s));
    HIP_OK(run_ens_bc3(launch(e, EK_BC3), 1, pgrid(e, EK_BC3), s));
    HIP_OK(run_ens_bulk_partial(launch(e, EK_BULK), 0, pgrid(e, EK_BULK), s));
    HIP_OK(run_ens_bulk_final(launch(e, EK_ONE), 0, pgrid(e, EK_ONE), s));
    HIP_OK(run_ens_bulk_partial(launch(e, EK_BULK), 1, pgrid(e, EK_BULK), s));
    HIP_OK(run_ens_bulk_final(launch(e, EK_ONE), 1, pgrid(e, EK_ONE), s));
    HIP_OK(run_ens_bcu(launch(e, EK_BCU), pgrid(e, EK_BCU), s));
    HIP_OK(run_ens_advance(launch(e, EK_ONE), s));
    return SLC_OK;
}

// order e->es after `stream`
int enter(slc_ens* e, void* stream) {
    HIP_OK(hipEventRecord(e->ev_in, static_cast<hipStream_t>(stream)));
    HIP_OK(hipStreamWaitEvent(e->es, e->ev_in, 0));
    return SLC_OK;
}

// order `stream` after e->es
int leave(slc_ens* e, void* stream) {
    HIP_OK(hipEventRecord(e->ev_out, e->es));
    HIP_OK(hipStreamWaitEvent(static_cast<hipStream_t>(stream), e->ev_out, 0));
    return SLC_OK;
}

int64_t member_bytes(int nx, int ny, int nz, int order, int storage) {
    int64_t b = 0;
    for (int c = 0; c < 3; ++c) b += (int64_t)(kFieldSets + 1) * 8 * field_elems(c, nx, ny, nz);
    b += 8 * p_elems(nx, ny, nz);
    b += 3L * 8 * nx * ny * nz;  // near feet
    if (storage == 32)
        for (int c = 0; c < 3; ++c) b += 2L * 4 * field_elems(c, nx, ny, nz);
    b += spectral_bytes(nx, ny, nz);
    // tables: z (2 P + 4 columns of nz+2), arrival, Poisson, z-solve factors,
    // bulk and diagnostics partials
    b += 8L * (nz + 2) * (2 * order + 8 + 19) + 8L * (nx + ny) * 4 + 4L * 3 * nz +
         8L * bulk_blocks(nx, ny, nz) + 8L * 5 * kDiagMaxBlocks + 4096;
    return b;
}

}  // namespace

extern "C" {

int slc_ens_create(int device, int order, int storage_bits, slc_ens** out) {
    if (!out) return fail(SLC_EINVAL, "null argument");
    *out = nullptr;
    if (order != 4 && order != 6) return fail(SLC_EINVAL, "order must be 4 or 6, got %d", order);
    if (storage_bits != 64 && storage_bits != 32)
        return fail(SLC_EINVAL, "storage must be 32 or 64 bits, got %d", storage_bits);
    DeviceGuard dg(device);
    slc_ens* e = new slc_ens();
    e->device = device;
    e->order = order;
    e->storage = storage_bits;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        ncu > 0)
        e->ncu = ncu;
    hipError_t err = hipStreamCreateWithFlags(&e->es, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&e->ev_in, hipEventDisableTiming);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&e->ev_out, hipEventDisableTiming);
    if (err != hipSuccess) {
        slc_ens_destroy(e);
        return fail(SLC_EHIP, "ensemble stream/events: %s", hipGetErrorString(err));
    }
    rocfft_setup();
    if (rocfft_execution_info_create(&e->info) != rocfft_status_success) {
        slc_ens_destroy(e);
        return fail(SLC_EFFT, "rocfft_execution_info_create failed");
    }
    *out = e;
    return SLC_OK;
}

int slc_ens_destroy(slc_ens* e) {
    if (!e) return SLC_OK;
    DeviceGuard dg(e->device);
    if (e->es) (void)hipStreamSynchronize(e->es);
    drop_graph(e);
    for (Member* m : e->members) free_member(m);
    for (Member* m : e->pool) free_member(m);
    for (auto& c : e->classes) free_class(c);
    for (auto& st : e->stage) {
        if (st.host) (void)hipHostFree(st.host);
        if (st.done) (void)hipEventDestroy(st.done);
    }
    if (e->d_mem) (void)hipFree(e->d_mem);
    if (e->d_start) (void)hipFree(e->d_start);
    for (int k = 0; k < 3; ++k)
        if (e->d_umap[k]) (void)hipFree(e->d_umap[k]);
    if (e->fft_work) (void)hipFree(e->fft_work);
    if (e->info) rocfft_execution_info_destroy(e->info);
    if (e->ev_in) (void)hipEventDestroy(e->ev_in);
    if (e->ev_out) (void)hipEventDestroy(e->ev_out);
    if (e->es) (void)hipStreamDestroy(e->es);
    delete e;
    return SLC_OK;
}

int slc_ens_member_bytes(const slc_grid_desc* d, int storage_bits, int64_t* bytes) {
    if (!d || !bytes) return fail(SLC_EINVAL, "null argument");
    if (d->nx < 1 || d->ny < 2 || d->nz < 2) return fail(SLC_EINVAL, "grid too small");
    *bytes = member_bytes(d->nx, d->ny, d->nz, d->order, storage_bits);
    return SLC_OK;
}

// A new member arena of geometry `key` (plan, field sets, tables); on failure
// nothing stays allocated and the status is returned (SLC_ENOMEM when an
// allocation failed).
static int new_member(slc_ens* e, const slc_grid_desc* d, const GeoKey& key, Member** out) {
    *out = nullptr;
    Member* m = new Member();
    m->key = key;
    int rc = plan_create(d, e->device, false, &m->plan);
    if (rc) {  // SLC_ENOMEM when a plan buffer did not fit
        delete m;
        return rc;
    }
    m->nx = d->nx;
    m->ny = d->ny;
    m->nz = d->nz;
    void* p = nullptr;
    const int nx = m->nx, ny = m->ny, nz = m->nz;
#define MALLOC(bytes, dst, T)                    \
    do {                                         \
        if ((rc = m_alloc(m, (bytes), &p))) {    \
            free_member(m);                      \
            return rc;                           \
        }                                        \
        dst = static_cast<T*>(p);                \
    } while (0)
    for (int s = 0; s < kFieldSets; ++s)
        for (int c = 0; c < 3; ++c)
            MALLOC(8 * field_elems(c, nx, ny, nz), m->fld[s][c], double);
    MALLOC(8 * p_elems(nx, ny, nz), m->p, double);
    if (e->storage == 32)
        for (int c = 0; c < 3; ++c) {
            MALLOC(4 * field_elems(c, nx, ny, nz), m->f32n[c], float);
            MALLOC(4 * field_elems(c, nx, ny, nz), m->f32m[c], float);
        }
    MALLOC(8 * 18 * (size_t)(nz + 2), m->ztab, double);
    MALLOC(8 * 4, m->scal, double);
    MALLOC(8 * ES_COUNT, m->stats, double);
    MALLOC(8, m->ncl, unsigned long long);
    MALLOC(sizeof(int), m->state, int);
#undef MALLOC
    *out = m;
    return SLC_OK;
}

int slc_ens_add(slc_ens* e, const slc_grid_desc* d, double dt, double U_bulk, int* id) {
    if (int rc = ens_check(e)) return rc;
    if (!d || !id) return fail(SLC_EINVAL, "null argument");
    if (d->order != e->order)
        return fail(SLC_EINVAL, "member order %d differs from the ensemble's %d", d->order,
                    e->order);
    if (!(dt > 0.0)) return fail(SLC_EINVAL, "dt must be positive");
    if (d->traj_order != 2)
        return fail(SLC_EINVAL, "ensemble members trace trilinear trajectories "
                                "(traj_interp_order 2), got %d", d->traj_order);
    if (d->nz < 3) return fail(SLC_EINVAL, "ensemble members need nz >= 3");
    if (zsolve_cols(0, d->nz) < 1) return fail(SLC_EINVAL, "nz=%d too large for the z-solve", d->nz);
    DeviceGuard dg(e->device);
    const GeoKey key = geo_key(d);
    Member* m = nullptr;
    int rc = SLC_OK;
    for (size_t i = 0; i < e->pool.size(); ++i)
        if (e->pool[i]->key == key) {  // reuse a released arena of this geometry
            m = e->pool[i];
            e->pool.erase(e->pool.begin() + i);
            break;
        }
    if (!m) {
        rc = new_member(e, d, key, &m);
        if (rc == SLC_ENOMEM && !e->pool.empty()) {
            // released arenas of other geometries still hold HBM: free them
            // (hipFree waits for the device) and try once more
            for (Member* q : e->pool) free_member(q);
            e->pool.clear();
            rc = new_member(e, d, key, &m);
        }
        if (rc) return rc;
    }
    m->plan->g.nu = d->nu;  // the only per-run grid parameter
    m->dt = dt;
    m->U_bulk = U_bulk;
    const int nx = m->nx, ny = m->ny, nz = m->nz;
    // zero fields, scratch, counters; state: a = 0, b = 1, c = 2, BDF1 next.
    // Stream-ordered on es: after any step still using a reused arena.
    hipStream_t s0 = e->es;
    hipError_t err = hipSuccess;
    for (int s = 0; s < kFieldSets && err == hipSuccess; ++s)
        for (int c = 0; c < 3 && err == hipSuccess; ++c)
            err = hipMemsetAsync(m->fld[s][c], 0, 8 * field_elems(c, nx, ny, nz), s0);
    if (err == hipSuccess) err = hipMemsetAsync(m->p, 0, 8 * p_elems(nx, ny, nz), s0);
    if (err == hipSuccess) err = hipMemsetAsync(m->scal, 0, 8 * 4, s0);
    if (err == hipSuccess) err = hipMemsetAsync(m->stats, 0, 8 * ES_COUNT, s0);
    if (err == hipSuccess) err = hipMemsetAsync(m->ncl, 0, 8, s0);
    m->host_state = ens_state(0, 1, 2, 0);
    if (err == hipSuccess) err = run_ens_set_int(m->state, m->host_state, s0);
    // z-solve factors: mode 0 (BDF1, dt), mode 1 (BDF2, 2dt/3); theta = 1
    for (int mode = 0; mode < 2 && err == hipSuccess; ++mode)
        for (int c = 0; c < 3 && err == hipSuccess; ++c)
            err = run_zfactor(m->plan->g, c, mode == 0 ? dt : (2.0 / 3.0) * dt, 1.0,
                              m->plan->d_dzc, m->plan->d_dzf,
                              m->ztab + (size_t)(mode * 9 + c * 3) * (nz + 2), s0);
    if (err != hipSuccess) {
        free_member(m);
        return fail(SLC_EHIP, "ensemble member init: %s", hipGetErrorString(err));
    }
    m->id = e->next_id++;
    e->members.push_back(m);
    e->dirty = true;
    *id = m->id;
    return SLC_OK;
}

int slc_ens_remove(slc_ens* e, int id) {
    if (int rc = ens_check(e)) return rc;
    DeviceGuard dg(e->device);
    for (size_t i = 0; i < e->members.size(); ++i)
        if (e->members[i]->id == id) {
            // keep the arena for a later member of the same geometry; the
            // oldest pooled arena is freed beyond pool_max (hipFree waits
            // for the device)
            e->pool.push_back(e->members[i]);
            e->members.erase(e->members.begin() + i);
            e->dirty = true;
            if (e->pool.size() > e->pool_max) {
                free_member(e->pool.front());
                e->pool.erase(e->pool.begin());
            }
            return SLC_OK;
        }
    return fail(SLC_EINVAL, "no member with id %d", id);
}

int slc_ens_count(const slc_ens* e, int* n) {
    if (int rc = ens_check(e)) return rc;
    if (!n) return fail(SLC_EINVAL, "null argument");
    *n = (int)e->members.size();
    return SLC_OK;
}

int slc_ens_bytes(const slc_ens* e, int64_t* bytes) {
    if (int rc = ens_check(e)) return rc;
    if (!bytes) return fail(SLC_EINVAL, "null argument");
    int64_t b = e->class_bytes + (int64_t)e->fft_work_bytes;
    for (const Member* m : e->members) b += m->bytes + m->plan->bytes;
    // released arenas kept for reuse still hold their HBM
    for (const Member* m : e->pool) b += m->bytes + m->plan->bytes;
    *bytes = b;
    return SLC_OK;
}

int slc_ens_trim_pool(slc_ens* e, int64_t keep_bytes, int64_t* pool_bytes) {
    if (int rc = ens_check(e)) return rc;
    DeviceGuard dg(e->device);
    auto held = [e]() {
        int64_t b = 0;
        for (const Member* m : e->pool) b += m->bytes + m->plan->bytes;
        return b;
    };
    if (keep_bytes >= 0)
        while (!e->pool.empty() && held() > keep_bytes) {
            free_member(e->pool.front());
            e->pool.erase(e->pool.begin());
        }
    if (pool_bytes) *pool_bytes = held();
    return SLC_OK;
}

int slc_ens_set_fields(slc_ens* e, int id, const double* u, const double* v, const double* w,
                       const double* p, void* stream) {
    if (int rc = ens_check(e)) return rc;
    Member* m = find(e, id);
    if (!m) return fail(SLC_EINVAL, "no member with id %d", id);
    if (!u || !v || !w) return fail(SLC_EINVAL, "null field");
    DeviceGuard dg(e->device);
    if (int rc = enter(e, stream)) return rc;
    const int a = ens_a(m->host_state);
    const double* src[3] = {u, v, w};
    for (int c = 0; c < 3; ++c)
        HIP_OK(hipMemcpyAsync(m->fld[a][c], src[c], 8 * field_elems(c, m->nx, m->ny, m->nz),
                              hipMemcpyDeviceToDevice, e->es));
    if (p)
        HIP_OK(hipMemcpyAsync(m->p, p, 8 * p_elems(m->nx, m->ny, m->nz), hipMemcpyDeviceToDevice,
                              e->es));
    else
        HIP_OK(hipMemsetAsync(m->p, 0, 8 * p_elems(m->nx, m->ny, m->nz), e->es));
    // restart semantics: next step is a BDF1 bootstrap from buffer set a
    m->host_state = ens_state(a, (a + 1) % 3, (a + 2) % 3, 0);
    HIP_OK(run_ens_set_int(m->state, m->host_state, e->es));
    return leave(e, stream);
}

int slc_ens_get_fields(slc_ens* e, int id, double* u, double* v, double* w, double* p,
                       void* stream) {
    if (int rc = ens_check(e)) return rc;
    Member* m = find(e, id);
    if (!m) return fail(SLC_EINVAL, "no member with id %d", id);
    DeviceGuard dg(e->device);
    if (int rc = enter(e, stream)) return rc;
    const int a = ens_a(m->host_state);
    double* dst[3] = {u, v, w};
    for (int c = 0; c < 3; ++c)
        if (dst[c])
            HIP_OK(hipMemcpyAsync(dst[c], m->fld[a][c], 8 * field_elems(c, m->nx, m->ny, m->nz),
                                  hipMemcpyDeviceToDevice, e->es));
    if (p)
        HIP_OK(hipMemcpyAsync(p, m->p, 8 * p_elems(m->nx, m->ny, m->nz), hipMemcpyDeviceToDevice,
                              e->es));
    return leave(e, stream);
}

int slc_ens_stats(slc_ens* e, int id, double* out, void* stream) {
    if (int rc = ens_check(e)) return rc;
    Member* m = find(e, id);
    if (!m) return fail(SLC_EINVAL, "no member with id %d", id);
    if (!out) return fail(SLC_EINVAL, "null argument");
    DeviceGuard dg(e->device);
    if (int rc = enter(e, stream)) return rc;
    HIP_OK(hipMemcpyAsync(out, m->stats, 8 * ES_COUNT, hipMemcpyDeviceToDevice, e->es));
    return leave(e, stream);
}

int slc_ens_diagnostics(slc_ens* e, int id, double* out, void* stream) {
    if (int rc = ens_check(e)) return rc;
    Member* m = find(e, id);
    if (!m) return fail(SLC_EINVAL, "no member with id %d", id);
    if (!out) return fail(SLC_EINVAL, "null argument");
    DeviceGuard dg(e->device);
    if (int rc = enter(e, stream)) return rc;
    const int a = ens_a(m->host_state);
    if (int rc = slc_diagnostics(m->plan, m->fld[a][0], m->fld[a][1], m->fld[a][2], m->p, m->dt,
                                 out, e->es))
        return rc;
    return leave(e, stream);
}

int slc_ens_step(slc_ens* e, void* stream) {
    if (int rc = ens_check(e)) return rc;
    if (e->members.empty()) return SLC_OK;
    DeviceGuard dg(e->device);
    if (e->dirty)
        if (int rc = rebuild(e)) return rc;
    if (int rc = enter(e, stream)) return rc;
    if (e->use_graph) {
        if (!e->gexec) {
            hipGraph_t g = nullptr;
            HIP_OK(hipStreamBeginCapture(e->es, hipStreamCaptureModeThreadLocal));
            int rc = enqueue_step(e);
            hipError_t err = hipStreamEndCapture(e->es, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (err != hipSuccess) return fail(SLC_EHIP, "step capture: %s", hipGetErrorString(err));
            err = hipGraphInstantiate(&e->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (err != hipSuccess) {
                e->gexec = nullptr;
                return fail(SLC_EHIP, "graph instantiate: %s", hipGetErrorString(err));
            }
        }
        HIP_OK(hipGraphLaunch(e->gexec, e->es));
    } else {
        if (int rc = enqueue_step(e)) return rc;
    }
    for (Member* m : e->members) m->host_state = ens_next(m->host_state);
    return leave(e, stream);
}

int slc_ens_set_option(slc_ens* e, int option, int value) {
    if (int rc = ens_check(e)) return rc;
    switch (option) {
        case SLC_ENS_OPT_TILE:
            if (value < 0 || value >= kEnsTileShapes)
                return fail(SLC_EINVAL, "tile shape %d out of range", value);
            if (value != e->tile_shape) {
                e->tile_shape = value;
                e->dirty = true;
            }
            return SLC_OK;
        case SLC_ENS_OPT_GRID:
            if (value < 0 || value > 32) return fail(SLC_EINVAL, "grid per CU %d out of range", value);
            if (value != e->grid_per_cu) {
                e->grid_per_cu = value;
                drop_graph(e);
            }
            return SLC_OK;
        case SLC_ENS_OPT_GRAPH:
            e->use_graph = value != 0;
            if (!e->use_graph) drop_graph(e);
            return SLC_OK;
        case SLC_ENS_OPT_BDF2_FUSED:
            if (value != 0 && value != 1) return fail(SLC_EINVAL, "bdf2 fused must be 0 or 1");
            if (value != e->bdf2_fused) {
                e->bdf2_fused = value;
                e->dirty = true;  // the captured step holds the other passes
            }
            return SLC_OK;
        case SLC_ENS_OPT_POOL: {
            if (value < 0) return fail(SLC_EINVAL, "pool size %d must be >= 0", value);
            DeviceGuard dg(e->device);
            e->pool_max = (size_t)value;
            while (e->pool.size() > e->pool_max) {
                free_member(e->pool.front());
                e->pool.erase(e->pool.begin());
            }
            return SLC_OK;
        }
        default:
            return fail(SLC_EINVAL, "unknown ensemble option %d", option);
    }
}

int slc_ens_get_option(const slc_ens* e, int option, int* value) {
    if (int rc = ens_check(e)) return rc;
    if (!value) return fail(SLC_EINVAL, "null argument");
    switch (option) {
        case SLC_ENS_OPT_TILE: *value = e->tile_shape; return SLC_OK;
        case SLC_ENS_OPT_GRID: *value = e->grid_per_cu; return SLC_OK;
        case SLC_ENS_OPT_GRAPH: *value = e->use_graph ? 1 : 0; return SLC_OK;
        case SLC_ENS_OPT_BDF2_FUSED: *value = e->bdf2_fused; return SLC_OK;
        case SLC_ENS_OPT_POOL: *value = (int)e->pool_max; return SLC_OK;
        default: return fail(SLC_EINVAL, "unknown ensemble option %d", option);
    }
}

// The timing sweeps of slc_ens_autotune on e->es; leaves the best tile shape
// and grid mode in e (rebuilt).  Any failure returns its status with e in
// whatever state the failing sweep left; the caller restores it.
static int autotune_sweeps(slc_ens* e, int reps, hipEvent_t t0, hipEvent_t t1,
                           std::string& rep) {
    int best = e->tile_shape;
    float best_ms = 1e30f;
    int rc = SLC_OK;
    for (int shape = 0; shape < kEnsTileShapes && rc == SLC_OK; ++shape) {
        e->tile_shape = shape;
        if ((rc = rebuild(e))) return rc;
        // U* (and the FP32 copies) so the far pass traces real trajectories
        if (hipError_t err = run_ens_ustar(launch(e, EK_ELEM), pgrid(e, EK_ELEM), e->es))
            return fail(SLC_EHIP, "autotune ustar: %s", hipGetErrorString(err));
        if ((rc = sl_passes(e))) return rc;  // warm-up
        (void)hipEventRecord(t0, e->es);
        for (int r = 0; r < reps && rc == SLC_OK; ++r) rc = sl_passes(e);
        (void)hipEventRecord(t1, e->es);
        if (rc) return rc;
        if (hipEventSynchronize(t1) != hipSuccess) return fail(SLC_EHIP, "autotune timing failed");
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, t0, t1);
        ms /= (float)reps;
        char line[64];
        snprintf(line, sizeof line, "%d %.9g\n", shape, ms);
        rep += line;
        if (ms < best_ms) {
            best_ms = ms;
            best = shape;
        }
    }
    e->tile_shape = best;
    e->dirty = true;
    if ((rc = rebuild(e))) return rc;
    // grid of the other batched kernels: time the substeps that only write
    // scratch (ghost fill of V^n, U*, z-solves of the predictor buffer,
    // divergence planes, Thomas)
    int best_grid = e->grid_per_cu;
    float best_gms = 1e30f;
    const int grids[3] = {0, 4, 16};
    for (int gi = 0; gi < 3; ++gi) {
        e->grid_per_cu = grids[gi];
        auto substeps = [&]() -> int {
            hipStream_t s = e->es;
            HIP_OK(run_ens_bc3(launch(e, EK_BC3), 0, pgrid(e, EK_BC3), s));
            HIP_OK(run_ens_ustar(launch(e, EK_ELEM), pgrid(e, EK_ELEM), s));
            for (int c = 0; c < 3; ++c)
                HIP_OK(run_ens_zsolve(launch(e, EK_ZS0 + c), c, e->zs_lds[c],
                                      pgrid(e, EK_ZS0 + c), e->zs_short[c], s));
            HIP_OK(run_ens_div_planes(launch(e, EK_PLANE), pgrid(e, EK_PLANE), s));
            HIP_OK(run_ens_thomas(launch(e, EK_THOMAS), pgrid(e, EK_THOMAS), s));
            return SLC_OK;
        };
        if ((rc = substeps())) return rc;
        (void)hipEventRecord(t0, e->es);
        for (int r = 0; r < reps && rc == SLC_OK; ++r) rc = substeps();
        (void)hipEventRecord(t1, e->es);
        if (rc) return rc;
        if (hipEventSynchronize(t1) != hipSuccess) return fail(SLC_EHIP, "autotune timing failed");
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, t0, t1);
        ms /= (float)reps;
        char line[64];
        snprintf(line, sizeof line, "grid%d %.9g\n", grids[gi], ms);
        rep += line;
        if (ms < best_gms) {
            best_gms = ms;
            best_grid = grids[gi];
        }
    }
    e->grid_per_cu = best_grid;
    // the timed passes accumulated clamp counts: clear them
    HIP_OK(run_ens_clear_ncl(launch(e, EK_ONE), e->es));
    return SLC_OK;
}

int slc_ens_autotune(slc_ens* e, int reps, void* stream, char* report, int64_t len) {
    if (int rc = ens_check(e)) return rc;
    if (reps < 1) return fail(SLC_EINVAL, "reps must be >= 1");
    if (e->members.empty()) return fail(SLC_EINVAL, "autotune needs members");
    DeviceGuard dg(e->device);
    if (int rc = enter(e, stream)) return rc;
    const int prev_shape = e->tile_shape, prev_grid = e->grid_per_cu;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    std::string rep;
    int rc = SLC_OK;
    if (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess)
        rc = fail(SLC_EHIP, "autotune: hipEventCreate failed");
    if (rc == SLC_OK) rc = autotune_sweeps(e, reps, t0, t1, rep);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (rc) {
        // keep the configuration the ensemble had; the launch tables are
        // rebuilt from it before the next step
        e->tile_shape = prev_shape;
        e->grid_per_cu = prev_grid;
        e->dirty = true;
        (void)leave(e, stream);
        return rc;
    }
    if (report && len > 0) snprintf(report, (size_t)len, "%s", rep.c_str());
    return leave(e, stream);
}

}  // extern "C"
