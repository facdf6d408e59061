// slc_ens_tile.h -- the batched SL tile kernel of the ensemble step:
// PASS 0 far pass (BDF2 members), PASS 1 near / bootstrap pass (BDF2 / BDF1
// members), PASS 2 fused pass (MODE_BDF2 / BDF1), each unit one sl_tile_body
// tile of its member.  Included once per
// (storage type, order) by slc_ens_tile_{d4,d6,f4,f6}.hip.
#pragma once

#include "slc_ens_dev.h"
#include "slc_tile.h"

namespace slc {

// One workgroup per tile: unit = xcd-ordered block index, its member from the
// per-unit member map (the persistent chunk walk of the other batched kernels
// keeps ~25 more registers live here and halves the occupancy).
template <typename T, int C, int P, int PASS, int TI, int TJ, int TK, int kEnsLds, int WAVES>
__global__ __launch_bounds__(kSB) __attribute__((amdgpu_waves_per_eu(WAVES))) void
ens_tile_kernel(const EnsMember* __restrict__ mem,
                                                       const unsigned* __restrict__ start,
                                                       const int* __restrict__ umap) {
    __shared__ __align__(16) T lds[kEnsLds / sizeof(T)];  // 16-byte pairs land here (stage_box)
    __shared__ int sh[PASS == 2 ? 12 : 6];
    __shared__ int s_cnt;
    __shared__ FootSt nst[PASS == 2 ? TI * TJ * TK : 1];
    const unsigned uu = (unsigned)xcd_block(gridDim.x);
    // member and state are the same for the whole workgroup: broadcast them
    // so the member table is read with scalar loads and its pointers stay in
    // scalar registers
    const int m = __builtin_amdgcn_readfirstlane(umap[uu]);
    const unsigned u = uu - start[m];
    const EnsMember& M = mem[m];
    const int st = __builtin_amdgcn_readfirstlane(*M.state);
    const int a = ens_a(st), b = ens_b(st), c = ens_c(st);
    const long NA = (long)M.g.nx * M.g.ny * M.g.nz;
    const Feet ft{M.nfoot, M.nfoot + NA, M.nfoot + 2 * NA};
    const double* Fa = M.fld[a][C];
    const double* Fb = M.fld[b][C];
    const T* Sa;
    const T* Sb;
    if constexpr (sizeof(T) == 4) {
        Sa = reinterpret_cast<const T*>(M.f32n[C]);
        Sb = reinterpret_cast<const T*>(M.f32m[C]);
    } else {
        Sa = reinterpret_cast<const T*>(Fa);
        Sb = reinterpret_cast<const T*>(Fb);
    }
    if (ens_bdf2(st)) {
        const TrajVel V{{M.us[0], M.us[1], M.us[2]}, {nullptr, nullptr, nullptr}};
        if (PASS == 0)
            sl_tile_body<T, C, P, MODE_FAR, TI, TJ, TK, kEnsLds>(
                M.g, M.at[C], V, Fa, Fb, Sa, Sb, M.dt, M.fld[c][C], ft, M.ncl, u, lds, sh, &s_cnt);
        else if (PASS == 2)
            sl_tile_body<T, C, P, MODE_BDF2, TI, TJ, TK, kEnsLds>(
                M.g, M.at[C], V, Fa, Fb, Sa, Sb, M.dt, M.fld[c][C], ft, M.ncl, u, lds, sh, &s_cnt,
                nst);
        else
            sl_tile_body<T, C, P, MODE_NEAR, TI, TJ, TK, kEnsLds>(
                M.g, M.at[C], V, Fa, Fb, Sa, Sb, M.dt, M.fld[c][C], ft, M.ncl, u, lds, sh, &s_cnt);
    } else if (PASS >= 1) {
        const TrajVel V{{M.fld[a][0], M.fld[a][1], M.fld[a][2]}, {nullptr, nullptr, nullptr}};
        sl_tile_body<T, C, P, MODE_BDF1, TI, TJ, TK, kEnsLds>(
            M.g, M.at[C], V, Fa, nullptr, Sa, nullptr, M.dt, M.fld[c][C], ft, M.ncl, u, lds, sh,
            &s_cnt);
    }
}

template <typename T, int C, int P, int PASS>
inline void launch_tile_shape(const EnsLaunch& L, int shape, unsigned grid, hipStream_t s) {
    switch (shape) {
        case 1:
            hipLaunchKernelGGL((ens_tile_kernel<T, C, P, PASS, 4, 4, 32, 40960, 1>), dim3(grid),
                               dim3(kSB), 0, s, L.mem, L.start, L.umap);
            break;
        case 2:
            hipLaunchKernelGGL((ens_tile_kernel<T, C, P, PASS, 2, 8, 16, 24576, 4>), dim3(grid),
                               dim3(kSB), 0, s, L.mem, L.start, L.umap);
            break;
        default:
            hipLaunchKernelGGL((ens_tile_kernel<T, C, P, PASS, 4, 4, 16, 24576, 4>), dim3(grid),
                               dim3(kSB), 0, s, L.mem, L.start, L.umap);
    }
}

template <typename T, int C, int P>
inline void launch_tile_pass(const EnsLaunch& L, int pass, int shape, unsigned grid, hipStream_t s) {
    if (pass == 0) launch_tile_shape<T, C, P, 0>(L, shape, grid, s);
    else if (pass == 1) launch_tile_shape<T, C, P, 1>(L, shape, grid, s);
    else launch_tile_shape<T, C, P, 2>(L, shape, grid, s);
}

template <typename T, int P>
inline void launch_tile_comp(const EnsLaunch& L, int comp, int pass, int shape, unsigned grid,
                             hipStream_t s) {
    if (comp == 0) launch_tile_pass<T, 0, P>(L, pass, shape, grid, s);
    else if (comp == 1) launch_tile_pass<T, 1, P>(L, pass, shape, grid, s);
    else launch_tile_pass<T, 2, P>(L, pass, shape, grid, s);
}

}  // namespace slc
