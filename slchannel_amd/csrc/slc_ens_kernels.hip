// slc_ens_kernels.hip -- batched persistent kernels of the ensemble step.
//
// Each kernel walks the concatenated work units of one kind (slc_ens.h) in
// contiguous chunks: logical workgroup l (XCD-aware order, so neighbouring
// chunks share an XCD's L2) owns units [l*T/G, (l+1)*T/G) and finds its first
// member by bisection of the prefix table, then steps through members as
// the units cross their boundaries.  The unit bodies are the single-channel
// ones.
#include "slc_ens_dev.h"
#include "slc_ops_dev.h"

namespace slc {


// ---------------------------------------------------------------------------
// ghost fills
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void ens_bc3_kernel(ENS_PARAMS, int role) {
    ens_for(start, nmem, total, [&](int m, unsigned u) {
        const EnsMember& M = mem[m];
        const int st = *M.state;
        const int set = role == 0 ? ens_a(st) : ens_c(st);
        const unsigned b0 = ublocks(M.nghost[0]), b1 = ublocks(M.nghost[1]);
        int c;
        unsigned lb;
        if (u < b0) { c = 0; lb = u; }
        else if (u < b0 + b1) { c = 1; lb = u - b0; }
        else { c = 2; lb = u - b0 - b1; }
        bc_body(M.g, M.fld[set][c], c, (long)lb * 256 + threadIdx.x);
    });
}

hipError_t run_ens_bc3(const EnsLaunch& L, int role, unsigned grid, hipStream_t s) {
    if (L.total == 0) return hipSuccess;
    hipLaunchKernelGGL(ens_bc3_kernel, dim3(grid), dim3(256), 0, s, ENS_ARGS(L), role);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void ens_bcp_kernel(ENS_PARAMS) {
    ens_for(start, nmem, total, [&](int m, unsigned u) {
        const EnsMember& M = mem[m];
        bc_body(M.g, M.p, 3, (long)u * 256 + threadIdx.x);
    });
}

hipError_t run_ens_bcp(const EnsLaunch& L, unsigned grid, hipStream_t s) {
    if (L.total == 0) return hipSuccess;
    hipLaunchKernelGGL(ens_bcp_kernel, dim3(grid), dim3(256), 0, s, ENS_ARGS(L));
    return hipGetLastError();
}

// u of the predictor set only (after the bulk shift, solver.py:720)
__global__ __launch_bounds__(256) void ens_bcu_kernel(ENS_PARAMS) {
    ens_for(start, nmem, total, [&](int m, unsigned u) {
        const EnsMember& M = mem[m];
        bc_body(M.g, M.fld[ens_c(*M.state)][0], 0, (long)u * 256 + threadIdx.x);
    });
}

hipError_t run_ens_bcu(const EnsLaunch& L, unsigned grid, hipStream_t s) {
    if (L.total == 0) return hipSuccess;
    hipLaunchKernelGGL(ens_bcu_kernel, dim3(grid), dim3(256), 0, s, ENS_ARGS(L));
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// U* (BDF2) or the V^{n-1} copy (BDF1), plus the FP32-storage copies
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void ens_ustar_kernel(ENS_PARAMS) {
    ens_for(start, nmem, total, [&](int m, unsigned u) {
        const EnsMember& M = mem[m];
        const int st = *M.state;
        const int a = ens_a(st), b = ens_b(st);
        const Grid& g = M.g;
        const long n[3] = {(long)(g.nx + 1) * (g.ny + 2) * (g.nz + 2),
                           (long)(g.nx + 2) * (g.ny + 1) * (g.nz + 2),
                           (long)(g.nx + 2) * (g.ny + 2) * (g.nz + 1)};
        const long i = (long)u * 256 + threadIdx.x;
        const bool f32 = M.f32n[0] != nullptr;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (i >= n[c]) continue;
            const double va = M.fld[a][c][i];
            if (ens_bdf2(st)) {
                const double vb = M.fld[b][c][i];
                M.us[c][i] = 2.0 * va - vb;
                if (f32) {
                    M.f32n[c][i] = (float)va;
                    M.f32m[c][i] = (float)vb;
                }
            } else {
                M.fld[b][c][i] = va;
                if (f32) M.f32n[c][i] = (float)va;
            }
        }
    });
}

hipError_t run_ens_ustar(const EnsLaunch& L, unsigned grid, hipStream_t s) {
    if (L.total == 0) return hipSuccess;
    hipLaunchKernelGGL(ens_ustar_kernel, dim3(grid), dim3(256), 0, s, ENS_ARGS(L));
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// SL tiles (slc_ens_tile.h, instantiated per storage type and order in
// slc_ens_tile_{d4,d6,f4,f6}.hip so they compile in parallel)
// ---------------------------------------------------------------------------

hipError_t run_ens_tile(const EnsLaunch& L, int comp, int pass, int order, int f32, int shape,
                        unsigned grid, hipStream_t s) {
    if (L.total == 0) return hipSuccess;
    if (f32) {
        if (order == 4) run_ens_tile_f4(L, comp, pass, shape, grid, s);
        else run_ens_tile_f6(L, comp, pass, shape, grid, s);
    } else {
        if (order == 4) run_ens_tile_d4(L, comp, pass, shape, grid, s);
        else run_ens_tile_d6(L, comp, pass, shape, grid, s);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// implicit z-diffusion (theta = 1) of the predictor, BDF1 or BDF2 factor
// ---------------------------------------------------------------------------

template <int C>
__global__ __launch_bounds__(256) void ens_zsolve_kernel(ENS_PARAMS) {
    extern __shared__ double lds[];
    __shared__ long cbase[kZColsMax];
    ens_for(start, nmem, total, [&](int m, unsigned u) {
        const EnsMember& M = mem[m];
        const int st = *M.state;
        const int nz = M.g.nz;
        const double* tab = M.ztab + (ens_bdf2(st) ? 9 : 0) * (long)(nz + 2) + C * 3L * (nz + 2);
        if (M.zch[C] > 0) {
            const long ncols = (long)M.g.nx * M.g.ny;
            const long col0 = ((long)u * kZScanCols + (threadIdx.x >> 6)) * zscan_k(M.zch[C]);
            if (col0 < ncols)
                zsolve_scan_any<C>(M.g, M.fld[ens_c(st)][C], tab, col0, ncols, M.zch[C]);
        } else {
            zsolve_body<C>(M.g, M.fld[ens_c(st)][C], tab, 0.0, M.dzc, M.dzf, M.zcols[C], u, lds,
                           cbase);
        }
    });
}

// The same for ensembles whose every member's columns hold at most 2
// unknowns per lane: only those variants are compiled in, so the kernel is
// built for more waves per SIMD than the general one (which the longest
// variants and the LDS column path hold at 125 VGPRs, four waves).
#ifndef SLC_ENS_ZS_SHORT_WAVES
#define SLC_ENS_ZS_SHORT_WAVES 6
#endif
template <int C>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SLC_ENS_ZS_SHORT_WAVES)))
void ens_zsolve_short_kernel(ENS_PARAMS) {
    static_assert(zscan_k(1) == zscan_k(2), "one column count for both short variants");
    ens_for(start, nmem, total, [&](int m, unsigned u) {
        const EnsMember& M = mem[m];
        const int st = *M.state;
        const int nz = M.g.nz;
        const double* tab = M.ztab + (ens_bdf2(st) ? 9 : 0) * (long)(nz + 2) + C * 3L * (nz + 2);
        const long ncols = (long)M.g.nx * M.g.ny;
        const long col0 = ((long)u * kZScanCols + (threadIdx.x >> 6)) * zscan_k(1);
        double* F = M.fld[ens_c(st)][C];
        if (col0 < ncols) {
            if (M.zch[C] == 1) zsolve_scan_cols<C, 1, zscan_k(1)>(M.g, F, tab, col0, ncols);
            else zsolve_scan_cols<C, 2, zscan_k(2)>(M.g, F, tab, col0, ncols);
        }
    });
}

hipError_t run_ens_zsolve(const EnsLaunch& L, int comp, size_t lds_bytes, unsigned grid,
                          bool short_cols, hipStream_t s) {
    if (L.total == 0) return hipSuccess;
#define SLC_ENS_ZS(K, CV) \
    hipLaunchKernelGGL((K<CV>), dim3(grid), dim3(256), lds_bytes, s, ENS_ARGS(L))
    if (short_cols) {
        if (comp == 0) SLC_ENS_ZS(ens_zsolve_short_kernel, 0);
        else if (comp == 1) SLC_ENS_ZS(ens_zsolve_short_kernel, 1);
        else SLC_ENS_ZS(ens_zsolve_short_kernel, 2);
    } else {
        if (comp == 0) SLC_ENS_ZS(ens_zsolve_kernel, 0);
        else if (comp == 1) SLC_ENS_ZS(ens_zsolve_kernel, 1);
        else SLC_ENS_ZS(ens_zsolve_kernel, 2);
    }
#undef SLC_ENS_ZS
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// projection: divergence -> planes, Thomas, gradient update
// ---------------------------------------------------------------------------

__device__ inline double ens_dt_eff(const EnsMember& M, int st) {
    return ens_bdf2(st) ? (2.0 / 3.0) * M.dt : M.dt;
}

__global__ __launch_bounds__(256) void ens_div_planes_kernel(ENS_PARAMS) {
    __shared__ double tile[kT][kT + 1];
    ens_for(start, nmem, total, [&](int m, unsigned u) {
        const EnsMember& M = mem[m];
        const int st = *M.state;
        const int c = ens_c(st);
        int i, j0, k0;
        plane_tile(M.g, u, i, j0, k0);
        div_planes_body(M.g, M.fld[c][0], M.fld[c][1], M.fld[c][2], M.dzf, ens_dt_eff(M, st),
                        M.planes, i, j0, k0, tile);
    });
}

hipError_t run_ens_div_planes(const EnsLaunch& L, unsigned grid, hipStream_t s) {
    if (L.total == 0) return hipSuccess;
    hipLaunchKernelGGL(ens_div_planes_kernel, dim3(grid), dim3(256), 0, s, ENS_ARGS(L));
    return hipGetLastError();
}

// one unit = one block of the single-run solve (the partition solve of
// spike_modes(nz) modes, or 256 modes of the sequential sweep), so members
// get the single run's doubles
__global__ __launch_bounds__(256) void ens_thomas_kernel(ENS_PARAMS) {
    __shared__ double lds[256 * 8];
    ens_for(start, nmem, total, [&](int m, unsigned u) {
        const EnsMember& M = mem[m];
        if (spike_ok(M.pt.nz))
            poisson_spike_body(M.pt, reinterpret_cast<double2*>(M.dhat), (long)u, lds);
        else
            poisson_thomas_body(M.pt, reinterpret_cast<double2*>(M.dhat), M.cpw,
                                (long)u * 256 + threadIdx.x);
    });
}

hipError_t run_ens_thomas(const EnsLaunch& L, unsigned grid, hipStream_t s) {
    if (L.total == 0) return hipSuccess;
    hipLaunchKernelGGL(ens_thomas_kernel, dim3(grid), dim3(256), 0, s, ENS_ARGS(L));
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void ens_project_kernel(ENS_PARAMS) {
    __shared__ double t0[kT + 1][kT + 2];
    __shared__ double t1[kT][kT + 1];
    ens_for(start, nmem, total, [&](int m, unsigned u) {
        const EnsMember& M = mem[m];
        const int st = *M.state;
        const int c = ens_c(st);
        int i, j0, k0;
        plane_tile(M.g, u, i, j0, k0);
        project_planes_body(M.g, M.fld[c][0], M.fld[c][1], M.fld[c][2], M.planes, M.inv_n,
                            ens_dt_eff(M, st), M.dzc, M.p, i, j0, k0, t0, t1);
    });
}

hipError_t run_ens_project(const EnsLaunch& L, unsigned grid, hipStream_t s) {
    if (L.total == 0) return hipSuccess;
    hipLaunchKernelGGL(ens_project_kernel, dim3(grid), dim3(256), 0, s, ENS_ARGS(L));
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// bulk constraint and the end-of-step state advance
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kRedBlock) void ens_bulk_partial_kernel(ENS_PARAMS, int shift) {
    __shared__ double red[kRedBlock];
    ens_for(start, nmem, total, [&](int m, unsigned u) {
        const EnsMember& M = mem[m];
        const int st = *M.state;
        bulk_partial_body(M.g, M.fld[ens_c(st)][0], M.cv, shift ? M.scal[0] : 0.0, shift,
                          M.partials, u, M.npart, red);
    });
}

hipError_t run_ens_bulk_partial(const EnsLaunch& L, int shift, unsigned grid, hipStream_t s) {
    if (L.total == 0) return hipSuccess;
    hipLaunchKernelGGL(ens_bulk_partial_kernel, dim3(grid), dim3(kRedBlock), 0, s, ENS_ARGS(L), shift);
    return hipGetLastError();
}

__global__ __launch_bounds__(kRedBlock) void ens_bulk_final_kernel(ENS_PARAMS, int mode) {
    __shared__ double red[kRedBlock];
    ens_for(start, nmem, total, [&](int m, unsigned) {
        const EnsMember& M = mem[m];
        bulk_finalize_body(M.partials, M.npart, M.total_volume, M.U_bulk, M.dt, mode, M.scal,
                           red);
    });
}

hipError_t run_ens_bulk_final(const EnsLaunch& L, int mode, unsigned grid, hipStream_t s) {
    if (L.total == 0) return hipSuccess;
    hipLaunchKernelGGL(ens_bulk_final_kernel, dim3(grid), dim3(kRedBlock), 0, s, ENS_ARGS(L), mode);
    return hipGetLastError();
}

// one thread per member: publish the step's statistics, clear the clamp
// counter, rotate the buffer sets
__global__ void ens_advance_kernel(const EnsMember* __restrict__ mem, int nmem) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmem) return;
    const EnsMember& M = mem[m];
    double* S = M.stats;
    S[ES_UBULK] = M.scal[2];
    S[ES_FORCING] = M.scal[1];
    S[ES_NCLAMP] = (double)*M.ncl;
    S[ES_STEPS] = S[ES_STEPS] + 1.0;
    S[ES_UBULK_PRE] = M.scal[3];
    *M.ncl = 0ull;
    *M.state = ens_next(*M.state);
}

hipError_t run_ens_advance(const EnsLaunch& L, hipStream_t s) {
    if (L.nmem == 0) return hipSuccess;
    hipLaunchKernelGGL(ens_advance_kernel, dim3((L.nmem + 63) / 64), dim3(64), 0, s, L.mem,
                       L.nmem);
    return hipGetLastError();
}

__global__ void ens_clear_ncl_kernel(const EnsMember* __restrict__ mem, int nmem) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < nmem) *mem[m].ncl = 0ull;
}

hipError_t run_ens_clear_ncl(const EnsLaunch& L, hipStream_t s) {
    if (L.nmem == 0) return hipSuccess;
    hipLaunchKernelGGL(ens_clear_ncl_kernel, dim3((L.nmem + 63) / 64), dim3(64), 0, s, L.mem,
                       L.nmem);
    return hipGetLastError();
}

__global__ void ens_set_int_kernel(int* p, int v) { *p = v; }

hipError_t run_ens_set_int(int* p, int v, hipStream_t s) {
    hipLaunchKernelGGL(ens_set_int_kernel, dim3(1), dim3(1), 0, s, p, v);
    return hipGetLastError();
}

}  // namespace slc
