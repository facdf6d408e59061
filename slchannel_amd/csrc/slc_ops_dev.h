// slc_ops_dev.h -- device bodies of the step's field operators, shared by the
// single-run kernels (slc_ops.hip) and the batched ensemble kernels
// (slc_ens_kernels.hip), so both compute the very same arithmetic:
//   ghost/boundary fill            solver.py:70-107, projection.py:180-195
//   theta-method z-diffusion       operators.py:382-630 (Thomas, shared factor)
//   divergence (+transpose)        utils.py:166-180
//   per-wavenumber Poisson Thomas  projection.py:133-247 (coefficients on the fly)
//   gradient projection            projection.py:250-272
//   bulk-flux constraint           solver.py:694-723, utils.py:147-163
// Each body takes its work coordinates explicitly (block or thread index
// within one channel) instead of reading blockIdx.
#pragma once

#include "slc_internal.h"

namespace slc {

// ---------------------------------------------------------------------------
// ghost fill
// ---------------------------------------------------------------------------
// The reference assigns ghost slabs in a fixed order (x, then y, then z), so
// every ghost's final value is +-(one interior value).  Each thread writes one
// ghost entry from its interior source; interiors are only read, so there is
// no ordering hazard.  Entries are enumerated as
//   S1: whole x-ghost planes          (contiguous)
//   S2: y-ghost rows of interior x    (contiguous along z)
//   S3: the two z ends of interior columns.
// Field kinds: 0 u, 1 v, 2 w, 3 p.

struct GhostShape {
    int NI, NJ, NK;
    int nxg, xg0, xg1;  // x-ghost planes
    int nyg, yg0, yg1;  // y-ghost rows
    long S1, S2, S3;
};

__host__ __device__ inline GhostShape ghost_shape(int kind, int nx, int ny, int nz) {
    GhostShape s;
    s.NI = (kind == 0) ? nx + 1 : nx + 2;
    s.NJ = (kind == 1) ? ny + 1 : ny + 2;
    s.NK = (kind == 2) ? nz + 1 : nz + 2;
    if (kind == 0) { s.nxg = 1; s.xg0 = 0; s.xg1 = 0; }
    else { s.nxg = 2; s.xg0 = 0; s.xg1 = nx + 1; }
    if (kind == 1) { s.nyg = 1; s.yg0 = 0; s.yg1 = 0; }
    else { s.nyg = 2; s.yg0 = 0; s.yg1 = ny + 1; }
    s.S1 = (long)s.nxg * s.NJ * s.NK;
    s.S2 = (long)nx * s.nyg * s.NK;
    s.S3 = (long)nx * ny * 2;
    return s;
}

__host__ __device__ inline long ghost_count(int kind, int nx, int ny, int nz) {
    const GhostShape s = ghost_shape(kind, nx, ny, nz);
    return s.S1 + s.S2 + s.S3;
}

__device__ inline void ghost_value(int kind, const Grid& g, const double* F, const GhostShape& s,
                                   int I, int J, int K, double* out_val) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    int is = I, js = J, ks = K;
    double sg = 1.0;
    bool zero = false;
    if (I == 0) is = nx;
    else if (kind != 0 && I == nx + 1) is = 1;
    if (J == 0) js = ny;
    else if (kind != 1 && J == ny + 1) js = 1;
    if (kind == 2) {
        if (K == 0 || K == nz) zero = true;
    } else if (kind == 3) {
        if (K == 0) ks = 1;
        else if (K == nz + 1) ks = nz;
    } else {
        if (K == 0) { ks = 1; sg = -1.0; }
        else if (K == nz + 1) { ks = nz; sg = g.top_neumann ? 1.0 : -1.0; }
    }
    double v = zero ? 0.0 : F[((long)is * s.NJ + js) * s.NK + ks];
    *out_val = (sg < 0.0) ? -v : v;
}

// Ghost entry t (0 <= t < ghost_count) of field F of the given kind.
__device__ inline void bc_body(const Grid& g, double* F, int kind, long t) {
    const GhostShape s = ghost_shape(kind, g.nx, g.ny, g.nz);
    int I, J, K;
    if (t < s.S1) {
        long pl = (long)s.NJ * s.NK;
        int p = (int)(t / pl);
        long r = t % pl;
        I = p == 0 ? s.xg0 : s.xg1;
        J = (int)(r / s.NK);
        K = (int)(r % s.NK);
    } else if (t < s.S1 + s.S2) {
        long t2 = t - s.S1;
        long row = t2 / s.NK;
        K = (int)(t2 % s.NK);
        I = 1 + (int)(row / s.nyg);
        J = (row % s.nyg) == 0 ? s.yg0 : s.yg1;
    } else if (t < s.S1 + s.S2 + s.S3) {
        long t3 = t - s.S1 - s.S2;
        long col = t3 >> 1;
        I = 1 + (int)(col / g.ny);
        J = 1 + (int)(col % g.ny);
        K = (t3 & 1) ? s.NK - 1 : 0;
    } else {
        return;
    }
    double v;
    ghost_value(kind, g, F, s, I, J, K, &v);
    F[((long)I * s.NJ + J) * s.NK + K] = v;
}

// ---------------------------------------------------------------------------
// implicit z-diffusion: (I - theta dt nu Dzz) x = d + (1-theta) dt nu Dzz d
// ---------------------------------------------------------------------------
// The matrix is the same for every (i, j) column, so it is factored once
// (tab = [a | c' | 1/den], n entries each) and every column only runs
// forward/back substitution.  Columns are staged through LDS so the HBM
// traffic is one coalesced read and one coalesced write per unknown.

constexpr int kZColsMax = 16;
constexpr int kZLdsBytes = 48 * 1024;  // dynamic LDS budget per workgroup

__host__ __device__ inline int zsolve_ld(int comp, int nz) {
    const int SZ = comp == 2 ? nz + 1 : nz + 2;
    return SZ + ((SZ & 1) ? 0 : 1);  // odd stride: conflict-free column walk
}

// columns per workgroup (0: a column does not fit the LDS budget)
__host__ __device__ inline int zsolve_cols(int comp, int nz) {
    int cols = kZLdsBytes / (zsolve_ld(comp, nz) * (int)sizeof(double));
    return cols > kZColsMax ? kZColsMax : cols;
}

// Column block cb (kZCols columns) of component C.  lds: kZCols * ld doubles;
// cbase: kZColsMax longs.  Ends with a barrier-free write-back; the caller
// synchronises before reusing lds.
template <int C>
__device__ inline void zsolve_body(const Grid& g, double* __restrict__ F,
                                   const double* __restrict__ tab, double beta,
                                   const double* __restrict__ dz_c,
                                   const double* __restrict__ dz_f, int kZCols, long cb,
                                   double* __restrict__ lds, long* cbase) {
    const int SZ = (int)Comp<C>::SZ(g);
    const long SY = Comp<C>::SY(g);
    const int ld = zsolve_ld(C, g.nz);
    const int n = (C == 2) ? g.nz - 1 : g.nz;
    const long ncols = (long)g.nx * g.ny;
    const long col0 = cb * kZCols;
    const int tid = threadIdx.x;

    // column base addresses (one 32-bit division per column, not per element)
    if (tid < kZCols) {
        const long gc = col0 + tid;
        long a = -1;
        if (gc < ncols) {
            const unsigned ug = (unsigned)gc;
            const int I = 1 + (int)(ug / (unsigned)g.ny), J = 1 + (int)(ug % (unsigned)g.ny);
            a = ((long)I * SY + J) * SZ;
        }
        cbase[tid] = a;
    }
    __syncthreads();
    // coalesced load of kZCols whole columns (ghost ends included)
    for (int e = tid; e < kZCols * SZ; e += blockDim.x) {
        const int c = e / SZ, k = e - (e / SZ) * SZ;
        const long a = cbase[c];
        if (a >= 0) lds[c * ld + k] = F[a + k];
    }
    __syncthreads();

    if (tid < kZCols && col0 + tid < ncols) {
        double* s = lds + tid * ld;
        if (beta != 0.0) {
            // explicit part, operators.py:446-466 (u, v) and 609-623 (w)
            if (C < 2) {
                double lo = s[0], mid = s[1];
                for (int k = 0; k < n; ++k) {
                    double hi = s[k + 2];
                    double dzl = dz_c[k], dzr = dz_c[k + 1], dzc = dz_f[k];
                    double d2;
                    if (k == 0) {
                        d2 = ((hi - mid) / dzr - 2.0 * mid / dzl) / dzc;
                    } else if (k == n - 1) {
                        if (g.top_neumann) d2 = (-(mid - lo) / dzl) / dzc;
                        else d2 = (-2.0 * mid / dzr - (mid - lo) / dzl) / dzc;
                    } else {
                        d2 = ((hi - mid) / dzr - (mid - lo) / dzl) / dzc;
                    }
                    s[k + 1] = mid + beta * d2;
                    lo = mid;
                    mid = hi;
                }
            } else {
                // the reference's w stencil is centred on face m (not m+1)
                double a = 0.0, b = s[0], c = s[1];
                for (int m = 0; m < n; ++m) {
                    double dzl = dz_f[m], dzr = dz_f[m + 1], dzc = dz_c[m + 1];
                    double d2;
                    if (m == 0) d2 = ((c - b) / dzr - b / dzl) / dzc;
                    else if (m == n - 1) d2 = (-b / dzr - (b - a) / dzl) / dzc;
                    else d2 = ((c - b) / dzr - (b - a) / dzl) / dzc;
                    double nxt = s[m + 2];
                    s[m + 1] = c + beta * d2;
                    a = b;
                    b = c;
                    c = nxt;
                }
            }
        }
        const double* A = tab;
        const double* CP = tab + n;
        const double* INV = tab + 2 * n;
        double dp = s[1] * INV[0];
        s[1] = dp;
        for (int k = 1; k < n; ++k) {
            dp = (s[k + 1] - A[k] * dp) * INV[k];
            s[k + 1] = dp;
        }
        double x = dp;
        for (int k = n - 2; k >= 0; --k) {
            x = s[k + 1] - CP[k] * x;
            s[k + 1] = x;
        }
    }
    __syncthreads();

    for (int e = tid; e < kZCols * SZ; e += blockDim.x) {
        const int c = e / SZ, k = e - (e / SZ) * SZ;
        const long a = cbase[c];
        if (a >= 0 && k >= 1 && k <= n) F[a + k] = lds[c * ld + k];
    }
}

// z-solve of the theta-method, one column per wavefront.  Lane
// l holds unknowns [l CH, (l+1) CH) in registers.  Forward and back
// substitution of the shared factorisation are first-order linear
// recurrences: each lane runs its stretch as an affine map of the value
// entering it (the same operations as the sequential sweep where that value
// is known: lane 0 forward, the lane holding the last unknown backward), the
// 64 maps are composed with a Kogge-Stone scan (6 shuffle steps), and each
// lane applies its entering value.  Every unknown is read and written once;
// a wavefront's loads and stores cover its column's n contiguous doubles.
// Differs from the sequential sweep by rounding only (tridiagonal solves
// carry the 1e-13 tolerance, not bit-exactness).
constexpr int kZScanMaxCh = 16;
constexpr int kZScanCols = 4;  // columns (wavefronts) per 256-thread workgroup

// unknowns per lane (power of two), 0 when the column is too long for the
// register path (n > 64 kZScanMaxCh)
// Columns per wavefront of the ensemble's register z-solve: short columns
// (at most 2 unknowns per lane, 128 levels) are solved four to a wavefront,
// so that a wavefront has more than one or two 512-byte rows of loads in
// flight (48 members of 64x48x64 .. 192x128x128: 712 -> 555 us per
// component).  The single-run z-solve keeps one column per wavefront.
#ifndef SLC_ZSCAN_SHORT_K
#define SLC_ZSCAN_SHORT_K 4
#endif
__host__ __device__ constexpr int zscan_k(int ch) { return ch <= 2 ? SLC_ZSCAN_SHORT_K : 1; }

__host__ __device__ inline int zsolve_scan_ch(int comp, int nz) {
    const int n = comp == 2 ? nz - 1 : nz;
    int ch = 1;
    while (ch * 64 < n) ch *= 2;
    return ch <= kZScanMaxCh ? ch : 0;
}

// Explicit part of the theta-method right-hand side, x_j + beta Dzz x_j, on
// the lane's unknowns (the formulas and operation order of zsolve_body, from
// operators.py:446-466 for u, v and 609-623 for w).  p[i] holds unknown
// j0 + i and, for j0 + i == n, the entry just above the last unknown (the
// top ghost of u, v; the top wall face of w); s0 is the bottom ghost / wall
// face.  Neighbours across lanes come by shuffles (all lanes take part).
template <int C, int CH>
__device__ inline void zsolve_explicit(const Grid& g, double* p, int j0, int n, double beta,
                                       const double* __restrict__ dz_c,
                                       const double* __restrict__ dz_f, double s0) {
    double ext[CH + 3];  // unknowns j0-2 .. j0+CH
    ext[1] = __shfl_up(p[CH - 1], 1);
    ext[0] = CH >= 2 ? __shfl_up(p[CH >= 2 ? CH - 2 : 0], 1) : __shfl_up(p[0], 2);
    ext[CH + 2] = __shfl_down(p[0], 1);
#pragma unroll
    for (int i = 0; i < CH; ++i) ext[i + 2] = p[i];
#pragma unroll
    for (int i = 0; i < CH; ++i) {
        const int j = j0 + i;
        if (j >= n) continue;
        double d2;
        if (C < 2) {
            const double lo = ext[i + 1], mid = ext[i + 2], hi = ext[i + 3];
            const double dzl = dz_c[j], dzr = dz_c[j + 1], dzc = dz_f[j];
            if (j == 0) {
                d2 = ((hi - mid) / dzr - 2.0 * mid / dzl) / dzc;
            } else if (j == n - 1) {
                if (g.top_neumann) d2 = (-(mid - lo) / dzl) / dzc;
                else d2 = (-2.0 * mid / dzr - (mid - lo) / dzl) / dzc;
            } else {
                d2 = ((hi - mid) / dzr - (mid - lo) / dzl) / dzc;
            }
            p[i] = mid + beta * d2;
        } else {
            // centred on face j (value b), stored to face j + 1 (value c)
            const double c = ext[i + 2];
            const double b = (j == 0) ? s0 : ext[i + 1];
            const double a = (j == 1) ? s0 : ext[i];
            const double dzl = dz_f[j], dzr = dz_f[j + 1], dzc = dz_c[j + 1];
            if (j == 0) d2 = ((c - b) / dzr - b / dzl) / dzc;
            else if (j == n - 1) d2 = (-b / dzr - (b - a) / dzl) / dzc;
            else d2 = ((c - b) / dzr - (b - a) / dzl) / dzc;
            p[i] = c + beta * d2;
        }
    }
}

// Orders the wavefront's LDS accesses (a wavefront executes them in order;
// this keeps the compiler from moving them across).
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS doubles per column of the transposed z-solve (zsolve_scan_cols): one
// pad after every CH unknowns, so that lane-contiguous reads of CH values
// spread over the banks.
template <int CH>
__host__ __device__ constexpr int zscan_lds(int K = 1) {
    return 64 * (CH + 1) * K;
}

// K columns col0 .. col0 + K - 1 (those below ncols) per wavefront: the
// loads of all K are issued before any is solved (more bytes in flight per
// wavefront; the factor tables are shared), then each column is solved
// exactly as a column alone.  With xlds (zscan_lds<CH>(K) doubles of the
// wavefront's own LDS) the column moves between HBM and the registers
// through LDS: wavefront-contiguous loads and stores (each load instruction
// touches 512 consecutive bytes) instead of each lane reading its CH
// consecutive unknowns (an instruction touches every cache line of the
// column).  Same unknowns in the same registers, so the same results.
template <int C, int CH, int K>
__device__ inline void zsolve_scan_cols(const Grid& g, double* __restrict__ F,
                                        const double* __restrict__ tab, long col0, long ncols,
                                        double beta = 0.0,
                                        const double* __restrict__ dz_c = nullptr,
                                        const double* __restrict__ dz_f = nullptr,
                                        double* __restrict__ xlds = nullptr) {
    const int lane = threadIdx.x & 63;
    const int n = (C == 2) ? g.nz - 1 : g.nz;
    const double* A = tab;
    const double* CP = tab + n;
    const double* INV = tab + 2 * n;
    const int j0 = lane * CH;
    double* Fc[K];
    bool ok[K];
    double p[K][CH], q[K][CH];
#pragma unroll
    for (int kc = 0; kc < K; ++kc) {
        const long col = col0 + kc;
        ok[kc] = col < ncols;
        const unsigned ug = (unsigned)(ok[kc] ? col : col0);
        const int I = 1 + (int)(ug / (unsigned)g.ny), J = 1 + (int)(ug % (unsigned)g.ny);
        Fc[kc] = F + ((long)I * Comp<C>::SY(g) + J) * Comp<C>::SZ(g) + 1;
        if (xlds) {
            double* X = xlds + kc * zscan_lds<CH>();
#pragma unroll
            for (int m = 0; m < CH; ++m) {
                const int e = 64 * m + lane;
                X[e + e / CH] = (ok[kc] && e <= n) ? Fc[kc][e] : 0.0;
            }
        } else {
#pragma unroll
            for (int i = 0; i < CH; ++i)
                p[kc][i] = (ok[kc] && j0 + i <= n) ? Fc[kc][j0 + i] : 0.0;
        }
    }
    if (xlds) {
        wave_sync();
#pragma unroll
        for (int kc = 0; kc < K; ++kc)
#pragma unroll
            for (int i = 0; i < CH; ++i) p[kc][i] = xlds[kc * zscan_lds<CH>() + lane * (CH + 1) + i];
        wave_sync();
    }
    if (beta != 0.0) {
#pragma unroll
        for (int kc = 0; kc < K; ++kc)
            zsolve_explicit<C, CH>(g, p[kc], j0, n, beta, dz_c, dz_f, ok[kc] ? Fc[kc][-1] : 0.0);
    }
    // forward: dp_j = p_j + q_j y with y = dp_{j0-1}
    double P[K], Q[K];
#pragma unroll
    for (int kc = 0; kc < K; ++kc) {
        double pr = 0.0, qr = 1.0;
#pragma unroll
        for (int i = 0; i < CH; ++i) {
            const int j = j0 + i;
            if (j < n) {
                if (j == 0) {
                    pr = p[kc][i] * INV[0];
                    qr = 0.0;
                } else {
                    const double a = A[j], r = INV[j];
                    pr = (p[kc][i] - a * pr) * r;
                    qr = -(a * qr) * r;
                }
            }
            p[kc][i] = pr;
            q[kc][i] = qr;
        }
        P[kc] = pr;
        Q[kc] = qr;
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
        for (int kc = 0; kc < K; ++kc) {
            const double Pp = __shfl_up(P[kc], off), Qp = __shfl_up(Q[kc], off);
            if (lane >= off) {
                P[kc] = P[kc] + Q[kc] * Pp;
                Q[kc] = Q[kc] * Qp;
            }
        }
    }
#pragma unroll
    for (int kc = 0; kc < K; ++kc) {
        double y = __shfl_up(P[kc], 1);
        if (lane == 0) y = 0.0;
#pragma unroll
        for (int i = 0; i < CH; ++i) p[kc][i] = p[kc][i] + q[kc][i] * y;
    }
    // backward: x_j = r_j + t_j z with z = x_{j0+CH}
    double R[K], T[K];
#pragma unroll
    for (int kc = 0; kc < K; ++kc) {
        double rr = 0.0, tr = 1.0;
#pragma unroll
        for (int i = CH - 1; i >= 0; --i) {
            const int j = j0 + i;
            if (j < n) {
                if (j == n - 1) {
                    rr = p[kc][i];
                    tr = 0.0;
                } else {
                    const double c = CP[j];
                    rr = p[kc][i] - c * rr;
                    tr = -(c * tr);
                }
            }
            p[kc][i] = rr;
            q[kc][i] = tr;
        }
        R[kc] = rr;
        T[kc] = tr;
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
        for (int kc = 0; kc < K; ++kc) {
            const double Rn = __shfl_down(R[kc], off), Tn = __shfl_down(T[kc], off);
            if (lane + off < 64) {
                R[kc] = R[kc] + T[kc] * Rn;
                T[kc] = T[kc] * Tn;
            }
        }
    }
#pragma unroll
    for (int kc = 0; kc < K; ++kc) {
        double z = __shfl_down(R[kc], 1);
        if (lane == 63) z = 0.0;
        if (xlds) {
            double* X = xlds + kc * zscan_lds<CH>();
#pragma unroll
            for (int i = 0; i < CH; ++i) X[lane * (CH + 1) + i] = p[kc][i] + q[kc][i] * z;
            continue;
        }
        if (!ok[kc]) continue;
#pragma unroll
        for (int i = 0; i < CH; ++i)
            if (j0 + i < n) Fc[kc][j0 + i] = p[kc][i] + q[kc][i] * z;
    }
    if (xlds) {
        wave_sync();
#pragma unroll
        for (int kc = 0; kc < K; ++kc) {
            if (!ok[kc]) continue;
            const double* X = xlds + kc * zscan_lds<CH>();
#pragma unroll
            for (int m = 0; m < CH; ++m) {
                const int e = 64 * m + lane;
                if (e < n) Fc[kc][e] = X[e + e / CH];
            }
        }
    }
}

// Runtime CH dispatch (ch from zsolve_scan_ch), theta = 1: the zscan_k(ch)
// columns col0 .. of one wavefront (those below ncols).
template <int C>
__device__ inline void zsolve_scan_any(const Grid& g, double* __restrict__ F,
                                       const double* __restrict__ tab, long col0, long ncols,
                                       int ch) {
    switch (ch) {
        case 1: zsolve_scan_cols<C, 1, zscan_k(1)>(g, F, tab, col0, ncols); break;
        case 2: zsolve_scan_cols<C, 2, zscan_k(2)>(g, F, tab, col0, ncols); break;
        case 4: zsolve_scan_cols<C, 4, zscan_k(4)>(g, F, tab, col0, ncols); break;
        case 8: zsolve_scan_cols<C, 8, zscan_k(8)>(g, F, tab, col0, ncols); break;
        default: zsolve_scan_cols<C, 16, zscan_k(16)>(g, F, tab, col0, ncols); break;
    }
}

// ---------------------------------------------------------------------------
// divergence -> plane-major (k, i, j) Poisson right-hand side
// ---------------------------------------------------------------------------

constexpr int kT = 32;  // transpose tile edge

// tiles of the (i; j, k) transposes: nx * ceil(ny/kT) * ceil(nz/kT)
__host__ __device__ inline long plane_tiles(int nx, int ny, int nz) {
    return (long)nx * ((ny + kT - 1) / kT) * ((nz + kT - 1) / kT);
}

// tile index -> (i, j0, k0), k-tiles fastest
__device__ inline void plane_tile(const Grid& g, long t, int& i, int& j0, int& k0) {
    const int ntk = (g.nz + kT - 1) / kT, ntj = (g.ny + kT - 1) / kT;
    k0 = (int)(t % ntk) * kT;
    const long r = t / ntk;
    j0 = (int)(r % ntj) * kT;
    i = (int)(r / ntj);
}

__device__ inline double divergence_at(const Grid& g, const double* __restrict__ u,
                                       const double* __restrict__ v,
                                       const double* __restrict__ w,
                                       const double* __restrict__ dz_f, int I, int J, int K) {
    const long uSY = g.ny + 2, uSZ = g.nz + 2;
    const long vSY = g.ny + 1, vSZ = g.nz + 2;
    const long wSY = g.ny + 2, wSZ = g.nz + 1;
    double du = (u[((long)I * uSY + J) * uSZ + K] - u[((long)(I - 1) * uSY + J) * uSZ + K]) / g.dx;
    double dv = (v[((long)I * vSY + J) * vSZ + K] - v[((long)I * vSY + (J - 1)) * vSZ + K]) / g.dy;
    long wb = ((long)I * wSY + J) * wSZ;
    double dw = (w[wb + K] - w[wb + K - 1]) / dz_f[K - 1];
    return du + dv + dw;
}

// tile: kT x (kT+1) doubles of LDS
__device__ inline void div_planes_body(const Grid& g, const double* __restrict__ u,
                                       const double* __restrict__ v,
                                       const double* __restrict__ w,
                                       const double* __restrict__ dz_f, double scale,
                                       double* __restrict__ planes, int i, int j0, int k0,
                                       double (*tile)[kT + 1]) {
    const int t = threadIdx.x;
    {
        const int kk = t % kT;
        for (int jj = t / kT; jj < kT; jj += 256 / kT) {
            int j = j0 + jj, k = k0 + kk;
            if (j < g.ny && k < g.nz) {
                double d = divergence_at(g, u, v, w, dz_f, i + 1, j + 1, k + 1);
                tile[jj][kk] = d / scale;
            }
        }
    }
    __syncthreads();
    {
        const int jj = t % kT;
        for (int kk = t / kT; kk < kT; kk += 256 / kT) {
            int j = j0 + jj, k = k0 + kk;
            if (j < g.ny && k < g.nz) planes[((long)k * g.nx + i) * g.ny + j] = tile[jj][kk];
        }
    }
}

// ---------------------------------------------------------------------------
// per-wavenumber Thomas (projection.py:199-247 with the matrices of 47-114)
// ---------------------------------------------------------------------------
// dhat: (nz, nx, nky) complex, plane-major; one lane per mode walks k, so a
// wavefront reads 64 consecutive modes (1 KiB) per plane.  c' is kept in a
// real scratch of the same (nz, modes) shape.

__device__ inline void poisson_thomas_body(const PoissonTabs& t, double2* __restrict__ d,
                                           double* __restrict__ cpw, long m) {
    const long M = (long)t.nx * t.nky;
    if (m >= M) return;
    const int nz = t.nz;
    const int i = (int)(m / t.nky), jc = (int)(m % t.nky);
    const double khsq = t.kxsq[i] + t.kysq[jc];
    // row 0 (Neumann bottom; identity row for the (0,0) mode)
    double b = -(t.cs[0] + khsq);
    b += t.cl[0];
    if (nz == 1) b += t.cr[0];
    double c = nz > 1 ? t.cr[0] : 0.0;
    if (m == 0) {
        b = 1.0;
        c = 0.0;
    }
    double cp = c / b;
    double2 dk = d[m];
    double dr = dk.x / b, di = dk.y / b;
    d[m] = make_double2(dr, di);
    cpw[m] = cp;
    for (int k = 1; k < nz; ++k) {
        const double a = t.cl[k];
        double bk = -(t.cs[k] + khsq);
        double ck = 0.0;
        if (k == nz - 1) bk += t.cr[nz - 1];
        else ck = t.cr[k];
        double den = bk - a * cp;
        if (k < nz - 1) cp = ck / den;
        long idx = (long)k * M + m;
        dk = d[idx];
        dr = (dk.x - a * dr) / den;
        di = (dk.y - a * di) / den;
        d[idx] = make_double2(dr, di);
        cpw[idx] = cp;
    }
    double xr = dr, xi = di;
    for (int k = nz - 2; k >= 0; --k) {
        long idx = (long)k * M + m;
        double cpk = cpw[idx];
        dk = d[idx];
        xr = dk.x - cpk * xr;
        xi = dk.y - cpk * xi;
        d[idx] = make_double2(xr, xi);
    }
}

// ---------------------------------------------------------------------------
// per-wavenumber tridiagonal solve by partition (the system of
// poisson_thomas_body, solved in another order: 32 B of memory traffic per
// complex unknown instead of 80)
// ---------------------------------------------------------------------------
// A 256-thread block takes spike_modes(nz) consecutive modes; its threads
// split each mode's nz levels into spike_segs(nz) segments of at most
// kSpikeR levels (thread t: mode t % MB, segment t / MB), so each plane row
// of the block is one contiguous run of MB complex values.  Per segment the
// thread solves, in registers, the segment's rows with the couplings to its
// neighbours removed, for the right-hand side (y) and for the two couplings
// (alpha: unit value one level below the segment, beta: one level above):
//   x = y + alpha L_prev + beta F_next,
// L_prev the last level of the segment below and F_next the first of the
// one above.  One thread per mode then solves the reduced system of the
// segments' first and last levels (a scalar two-term recurrence across the
// segments, forward then back), and every thread forms its levels.  Rows,
// the bottom Neumann row and the (0,0) mode's identity row are those of
// poisson_thomas_body (projection.py:47-114); the partition is exact
// algebra, so the solution agrees to rounding.  Needs nz <= 8 * 256; LDS:
// 256 x 8 doubles.
constexpr int kSpikeR = 8;

__host__ __device__ inline int spike_segs(int nz) {
    int s = 32;
    while (s * kSpikeR < nz) s *= 2;
    return s;
}

// modes per workgroup of NT threads
__host__ __device__ inline int spike_modes(int nz, int NT = 256) { return NT / spike_segs(nz); }

#ifndef SLC_SPIKE_MAX_NZ
#define SLC_SPIKE_MAX_NZ (kSpikeR * 256)
#endif
__host__ __device__ inline bool spike_ok(int nz) { return nz <= SLC_SPIKE_MAX_NZ; }

// row k of the Poisson system of mode m (khsq = kx^2 + ky^2): a (below),
// b (diagonal), c (above), as poisson_thomas_body builds them
__device__ inline void poisson_row(const PoissonTabs& t, long m, double khsq, int k, double& a,
                                   double& b, double& c) {
    const int nz = t.nz;
    if (k == 0) {
        a = 0.0;
        b = -(t.cs[0] + khsq);
        b += t.cl[0];
        if (nz == 1) b += t.cr[0];
        c = nz > 1 ? t.cr[0] : 0.0;
        if (m == 0) {
            b = 1.0;
            c = 0.0;
        }
        return;
    }
    a = t.cl[k];
    b = -(t.cs[k] + khsq);
    c = 0.0;
    if (k == nz - 1) b += t.cr[nz - 1];
    else c = t.cr[k];
}

// NT threads per workgroup (lds: NT * 8 doubles).
template <int NT = 256>
__device__ inline void poisson_spike_body(const PoissonTabs& t, double2* __restrict__ d,
                                          long unit, double* __restrict__ lds) {
    const int nz = t.nz;
    const long M = (long)t.nx * t.nky;
    const int S = spike_segs(nz), MB = NT / S;
    const int R = (nz + S - 1) / S;          // levels per segment (<= kSpikeR)
    const int nseg = (nz + R - 1) / R;       // segments holding levels
    const int ml = threadIdx.x % MB, seg = threadIdx.x / MB;
    const long m = unit * MB + ml;
    const bool live = m < M && seg < nseg;
    const int k0 = seg * R;
    const int nr = live ? min(R, nz - k0) : 0;
    double khsq = 0.0;
    if (live) khsq = t.kxsq[m / t.nky] + t.kysq[m % t.nky];
    double yr[kSpikeR], yi[kSpikeR], al[kSpikeR], be[kSpikeR], cp[kSpikeR];
    // forward elimination of the segment (couplings to the neighbours moved
    // to the alpha / beta right-hand sides)
#pragma unroll
    for (int i = 0; i < kSpikeR; ++i) {
        if (i < nr) {
            const int k = k0 + i;
            double a, b, c;
            poisson_row(t, m, khsq, k, a, b, c);
            const double2 dk = d[(long)k * M + m];
            const double ra = (i == 0) ? -a : 0.0;           // coupling below
            const double rb = (i == nr - 1) ? -c : 0.0;      // coupling above
            if (i == 0) {
                yr[0] = dk.x / b;
                yi[0] = dk.y / b;
                al[0] = ra / b;
                be[0] = rb / b;
                cp[0] = (nr > 1) ? c / b : 0.0;
            } else {
                const double den = b - a * cp[i - 1];
                yr[i] = (dk.x - a * yr[i - 1]) / den;
                yi[i] = (dk.y - a * yi[i - 1]) / den;
                al[i] = (0.0 - a * al[i - 1]) / den;
                be[i] = (rb - a * be[i - 1]) / den;
                cp[i] = (i < nr - 1) ? c / den : 0.0;
            }
        }
    }
#pragma unroll
    for (int i = kSpikeR - 2; i >= 0; --i) {
        if (i < nr - 1) {
            yr[i] -= cp[i] * yr[i + 1];
            yi[i] -= cp[i] * yi[i + 1];
            al[i] -= cp[i] * al[i + 1];
            be[i] -= cp[i] * be[i + 1];
        }
    }
    // the segment's first and last levels for the reduced system
    double* slot = lds + (seg * MB + ml) * 8;
    if (live) {
        slot[0] = yr[0];
        slot[1] = yi[0];
        slot[2] = al[0];
        slot[3] = be[0];
        slot[4] = yr[nr - 1];
        slot[5] = yi[nr - 1];
        slot[6] = al[nr - 1];
        slot[7] = be[nr - 1];
    }
    __syncthreads();
    // reduced system, one thread per mode: with L_{s-1} = P + Q F_s,
    // F_s = U_s + V_s F_{s+1} and L_s = P_s + Q_s F_{s+1}
    if (seg == 0 && m < M) {
        double Pr = 0.0, Pi = 0.0, Q = 0.0;
        for (int s = 0; s < nseg; ++s) {
            double* q = lds + (s * MB + ml) * 8;
            const double y0r = q[0], y0i = q[1], a0 = q[2], b0 = q[3];
            const double y1r = q[4], y1i = q[5], a1 = q[6], b1 = q[7];
            const double D = 1.0 - a0 * Q;
            const double Ur = (y0r + a0 * Pr) / D, Ui = (y0i + a0 * Pi) / D, V = b0 / D;
            const double nPr = y1r + a1 * Pr + a1 * Q * Ur, nPi = y1i + a1 * Pi + a1 * Q * Ui;
            const double nQ = a1 * Q * V + b1;
            q[0] = Ur;
            q[1] = Ui;
            q[2] = V;
            q[4] = nPr;
            q[5] = nPi;
            q[6] = nQ;
            Pr = nPr;
            Pi = nPi;
            Q = nQ;
        }
        double Fr = 0.0, Fi = 0.0;  // F_{s+1}
        for (int s = nseg - 1; s >= 0; --s) {
            double* q = lds + (s * MB + ml) * 8;
            const double Lr = q[4] + q[6] * Fr, Li = q[5] + q[6] * Fi;
            const double nFr = q[0] + q[2] * Fr, nFi = q[1] + q[2] * Fi;
            q[0] = nFr;  // F_s
            q[1] = nFi;
            q[4] = Lr;   // L_s
            q[5] = Li;
            Fr = nFr;
            Fi = nFi;
        }
    }
    __syncthreads();
    if (!live) return;
    double Lr = 0.0, Li = 0.0, Fr = 0.0, Fi = 0.0;
    if (seg > 0) {
        const double* q = lds + ((seg - 1) * MB + ml) * 8;
        Lr = q[4];
        Li = q[5];
    }
    if (seg < nseg - 1) {
        const double* q = lds + ((seg + 1) * MB + ml) * 8;
        Fr = q[0];
        Fi = q[1];
    }
#pragma unroll
    for (int i = 0; i < kSpikeR; ++i) {
        if (i < nr) {
            const double xr = yr[i] + al[i] * Lr + be[i] * Fr;
            const double xi = yi[i] + al[i] * Li + be[i] * Fi;
            d[(long)(k0 + i) * M + m] = make_double2(xr, xi);
        }
    }
}

// ---------------------------------------------------------------------------
// projection u -= dt grad p from the plane-major (unnormalised) inverse FFT
// ---------------------------------------------------------------------------
// Also writes the interior of the ghost-shaped pressure p = P / (nx ny).
// t0: (kT+1) x (kT+2), t1: kT x (kT+1) doubles of LDS.

__device__ inline void project_planes_body(const Grid& g, double* __restrict__ u,
                                           double* __restrict__ v, double* __restrict__ w,
                                           const double* __restrict__ P, double inv_n, double dt,
                                           const double* __restrict__ dz_c,
                                           double* __restrict__ p, int i, int j0, int k0,
                                           double (*t0)[kT + 2], double (*t1)[kT + 1]) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int i1 = (i + 1 == nx) ? 0 : i + 1;
    const int t = threadIdx.x;
    // load (j fastest: plane rows are contiguous in j)
    for (int e = t; e < (kT + 1) * (kT + 1); e += 256) {
        int jj = e % (kT + 1), kk = e / (kT + 1);
        int j = j0 + jj, k = k0 + kk;
        if (k < nz && j <= ny && jj <= kT) {
            int jw = (j == ny) ? 0 : j;
            t0[jj][kk] = P[((long)k * nx + i) * ny + jw];
        }
    }
    for (int e = t; e < kT * kT; e += 256) {
        int jj = e % kT, kk = e / kT;
        int j = j0 + jj, k = k0 + kk;
        if (k < nz && j < ny) t1[jj][kk] = P[((long)k * nx + i1) * ny + j];
    }
    __syncthreads();
    // update (k fastest: fields are z-contiguous)
    const int I = i + 1;
    const int kk = t % kT;
    for (int jj = t / kT; jj < kT; jj += 256 / kT) {
        int j = j0 + jj, k = k0 + kk;
        if (j >= ny || k >= nz) continue;
        const int J = j + 1, K = k + 1;
        double pc = t0[jj][kk] * inv_n;
        double pxp = t1[jj][kk] * inv_n;
        double pyp = t0[jj + 1][kk] * inv_n;
        if (u) {
            long ui = ((long)I * (ny + 2) + J) * (nz + 2) + K;
            u[ui] = u[ui] - dt * ((pxp - pc) / g.dx);
            long vi = ((long)I * (ny + 1) + J) * (nz + 2) + K;
            v[vi] = v[vi] - dt * ((pyp - pc) / g.dy);
        }
        if (u && k < nz - 1) {
            double pzp = t0[jj][kk + 1] * inv_n;
            long wi = ((long)I * (ny + 2) + J) * (nz + 1) + K;
            w[wi] = w[wi] - dt * ((pzp - pc) / dz_c[K]);
        }
        if (p) p[((long)I * (ny + 2) + J) * (nz + 2) + K] = pc;
    }
}

// ---------------------------------------------------------------------------
// bulk velocity and the exact flux shift
// ---------------------------------------------------------------------------
// sum_{ijk} u_ijk * (dx dy dz_f[k]) / V, reduced over a fixed partition and a
// fixed tree (deterministic).  scal: [0] correction, [1] forcing, [2] u_bulk
// after the shift, [3] u_bulk before it.

constexpr int kRedBlock = 256;

__host__ __device__ inline int bulk_blocks(int nx, int ny, int nz) {
    long N = (long)nx * ny * nz;
    long nb = (N + kRedBlock * 8 - 1) / (kRedBlock * 8);
    return (int)(nb < 2048 ? nb : 2048);
}

// red: kRedBlock doubles of LDS; every thread gets the block sum
__device__ inline double block_sum(double v, double* red) {
    __syncthreads();
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kRedBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    return red[0];
}

// Partial sum number blk of nb over u*cv; with shift, u += corr first.
__device__ inline void bulk_partial_body(const Grid& g, double* __restrict__ u,
                                         const double* __restrict__ cv, double corr, int shift,
                                         double* __restrict__ partials, long blk, long nb,
                                         double* red) {
    const long N = (long)g.nx * g.ny * g.nz;
    const long chunk = (N + nb - 1) / nb;
    const long beg = blk * chunk;
    const long end = beg + chunk < N ? beg + chunk : N;
    double acc = 0.0;
    long n = beg + threadIdx.x;
    if (n < end) {
        // (i, j, k) of the first element, then steps of kRedBlock elements in
        // the mixed radix (ny, nz): dk < nz and dj < ny, so one carry per
        // digit suffices.  Same elements in the same order as a division per
        // element.
        const unsigned un = (unsigned)n;
        int k = (int)(un % (unsigned)g.nz);
        const unsigned r = un / (unsigned)g.nz;
        int j = (int)(r % (unsigned)g.ny), i = (int)(r / (unsigned)g.ny);
        const int dk = kRedBlock % g.nz, t = kRedBlock / g.nz;
        const int dj = t % g.ny, di = t / g.ny;
#pragma unroll 4
        for (; n < end; n += kRedBlock) {
            const long ui = ((long)(i + 1) * (g.ny + 2) + (j + 1)) * (g.nz + 2) + (k + 1);
            double val = u[ui];
            if (shift) {
                val = val + corr;
                u[ui] = val;
            }
            acc += val * cv[k];
            k += dk;
            if (k >= g.nz) {
                k -= g.nz;
                ++j;
            }
            j += dj;
            if (j >= g.ny) {
                j -= g.ny;
                ++i;
            }
            i += di;
        }
    }
    double s = block_sum(acc, red);
    if (threadIdx.x == 0) partials[blk] = s;
}

// mode 0: scal[3] = ub*, scal[0] = U_b - ub*, scal[1] = scal[0] / dt
// mode 1: scal[2] = ub
__device__ inline void bulk_finalize_body(const double* __restrict__ partials, int np,
                                          double total_volume, double U_bulk, double dt,
                                          int mode, double* __restrict__ scal, double* red) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < np; b += kRedBlock) acc += partials[b];
    double s = block_sum(acc, red);
    if (threadIdx.x == 0) {
        double ub = s / total_volume;
        if (mode == 0) {
            double corr = U_bulk - ub;
            scal[3] = ub;
            scal[0] = corr;
            scal[1] = corr / dt;
        } else {
            scal[2] = ub;
        }
    }
}

}  // namespace slc
