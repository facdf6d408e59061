// slc_ens.h -- batched ensemble of independent channels: the per-member
// device table and the batched kernel launchers (slc_ens_kernels.hip).
//
// Every member is a complete SL-BDF2 channel of its own size, Reynolds number
// and time step.  One batched launch runs one substep of the step for ALL
// members: the work units of a kind (ghost-fill blocks, SL tiles, z-solve
// column blocks, transpose tiles, Thomas mode blocks, reduction partials)
// are concatenated member after member, and a persistent grid walks them in
// contiguous, XCD-aware chunks.  Each unit runs the very device body the
// single-channel kernels run (slc_tile.h, slc_ops_dev.h), so a member's
// result does not depend on who else is in the ensemble.
//
// Per-member state lives on the device and is advanced by the last kernel
// of the step, so a captured step graph replays without host updates:
//   bits 0-1  a   buffer set holding V^n
//   bits 2-3  b   buffer set holding V^{n-1}
//   bits 4-5  c   buffer set receiving the predictor (scratch)
//   bit  6        0: next step is the BDF1 bootstrap, 1: BDF2
// A BDF2 step rotates (a, b, c) -> (c, a, b) (solver.py:884-890); the BDF1
// bootstrap copies V^n into b and rotates (a, b, c) -> (c, b, a)
// (solver.py:733-765).
#pragma once

#include "slc_internal.h"

namespace slc {

__host__ __device__ inline int ens_a(int s) { return s & 3; }
__host__ __device__ inline int ens_b(int s) { return (s >> 2) & 3; }
__host__ __device__ inline int ens_c(int s) { return (s >> 4) & 3; }
__host__ __device__ inline bool ens_bdf2(int s) { return ((s >> 6) & 1) != 0; }
__host__ __device__ inline int ens_state(int a, int b, int c, int bdf2) {
    return a | (b << 2) | (c << 4) | (bdf2 << 6);
}
// state after one step from state s
__host__ __device__ inline int ens_next(int s) {
    return ens_bdf2(s) ? ens_state(ens_c(s), ens_a(s), ens_b(s), 1)
                       : ens_state(ens_c(s), ens_b(s), ens_a(s), 1);
}

// 256-wide blocks covering n entries
__host__ __device__ inline unsigned ublocks(long n) { return (unsigned)((n + 255) / 256); }

// per-member statistics published by the last kernel of every step
enum EnsStat {
    ES_UBULK = 0,   // bulk velocity after the shift (== U_bulk to rounding)
    ES_FORCING,     // correction / dt
    ES_NCLAMP,      // departure feet clamped in z during the step
    ES_STEPS,       // steps completed
    ES_UBULK_PRE,   // bulk velocity before the shift
    ES_COUNT
};

struct EnsMember {
    Grid g;
    ArrivalTabs at[3];
    double* fld[3][3];  // [buffer set][component]: ghost-shaped u, v, w
    double* p;          // ghost-shaped pressure
    double* us[3];      // U* = 2 V^n - V^{n-1}
    double* nfoot;      // near feet of one component (3 x arrival count)
    float* f32n[3];     // FP32-storage copies of V^n (null: FP64 gathers)
    float* f32m[3];     // ... and of V^{n-1}
    const double* dzc;
    const double* dzf;
    const double* cv;
    const double* ztab;  // [0: BDF1 (dt), 1: BDF2 (2dt/3)][comp] x 3 (nz+2)
    PoissonTabs pt;
    double* planes;  // this member's slot in its size class (nz, nx, ny)
    double* dhat;    // (nz, nx, nky) complex
    double* cpw;     // (nz, nx, nky)
    double* partials;
    double* scal;    // bulk scratch: [0] corr [1] forcing [2] ub after [3] ub before
    double* stats;   // EnsStat
    unsigned long long* ncl;
    int* state;
    double dt, U_bulk, total_volume, inv_n;
    int npart;
    int zcols[3];  // columns per z-solve work unit
    int zch[3];    // register z-solve unknowns per lane (0: LDS column path)
    long nghost[4];  // ghost entries of u, v, w, p
};

// Work-unit kinds; start[kind * (nmem + 1) + m] is the first unit of member m.
enum EnsKind {
    EK_BC3 = 0,  // ghost fill of u, v, w (256-entry blocks, u then v then w)
    EK_BCP,      // ghost fill of p
    EK_BCU,      // ghost fill of u only
    EK_ELEM,     // 256-element blocks over the largest velocity component
    EK_TILE0,    // SL tiles of u, v, w
    EK_TILE1,
    EK_TILE2,
    EK_ZS0,      // z-solve column blocks of u, v, w
    EK_ZS1,
    EK_ZS2,
    EK_PLANE,    // 32 x 32 transpose tiles (divergence, projection)
    EK_THOMAS,   // 256-mode Poisson Thomas blocks
    EK_BULK,     // bulk-reduction partials
    EK_ONE,      // one unit per member
    EK_COUNT
};

struct EnsLaunch {
    const EnsMember* mem;
    const unsigned* start;  // this kind's nmem + 1 prefix
    int nmem;
    unsigned total;
    const int* umap;        // member of every unit (SL tile kinds only)
};

// SL tile shapes the ensemble kernels are compiled for (autotuned):
//   0: 4 x 4 x 16, 24 KB LDS   1: 4 x 4 x 32, 40 KB LDS   2: 2 x 8 x 16, 24 KB LDS
//   (TI x TJ x TK arrival points per 256-thread workgroup)
constexpr int kEnsTileShapes = 3;
__host__ __device__ inline void ens_tile_shape(int s, int& ti, int& tj, int& tk) {
    ti = s == 2 ? 2 : 4;
    tj = s == 2 ? 8 : 4;
    tk = s == 1 ? 32 : 16;
}

// launchers: `grid` is the persistent grid size (workgroups); the SL tile
// launcher runs one workgroup per unit and ignores it
hipError_t run_ens_bc3(const EnsLaunch& L, int role, unsigned grid, hipStream_t s);
hipError_t run_ens_bcp(const EnsLaunch& L, unsigned grid, hipStream_t s);
hipError_t run_ens_bcu(const EnsLaunch& L, unsigned grid, hipStream_t s);
hipError_t run_ens_ustar(const EnsLaunch& L, unsigned grid, hipStream_t s);
// pass 0: far (BDF2 members), 1: near (BDF2) / bootstrap (BDF1)
hipError_t run_ens_tile(const EnsLaunch& L, int comp, int pass, int order, int f32, int shape,
                        unsigned grid, hipStream_t s);
void run_ens_tile_d4(const EnsLaunch& L, int comp, int pass, int shape, unsigned grid,
                     hipStream_t s);
void run_ens_tile_d6(const EnsLaunch& L, int comp, int pass, int shape, unsigned grid,
                     hipStream_t s);
void run_ens_tile_f4(const EnsLaunch& L, int comp, int pass, int shape, unsigned grid,
                     hipStream_t s);
void run_ens_tile_f6(const EnsLaunch& L, int comp, int pass, int shape, unsigned grid,
                     hipStream_t s);
hipError_t run_ens_zsolve(const EnsLaunch& L, int comp, size_t lds_bytes, unsigned grid,
                          bool short_cols, hipStream_t s);
hipError_t run_ens_div_planes(const EnsLaunch& L, unsigned grid, hipStream_t s);
hipError_t run_ens_thomas(const EnsLaunch& L, unsigned grid, hipStream_t s);
hipError_t run_ens_project(const EnsLaunch& L, unsigned grid, hipStream_t s);
hipError_t run_ens_bulk_partial(const EnsLaunch& L, int shift, unsigned grid, hipStream_t s);
hipError_t run_ens_bulk_final(const EnsLaunch& L, int mode, unsigned grid, hipStream_t s);
hipError_t run_ens_advance(const EnsLaunch& L, hipStream_t s);
hipError_t run_ens_clear_ncl(const EnsLaunch& L, hipStream_t s);
// *p = v, stream-ordered (no host memory involved)
hipError_t run_ens_set_int(int* p, int v, hipStream_t s);

}  // namespace slc
