// slc_device.h -- device-side grid description and the semi-Lagrangian
// building blocks shared by the gfx950 kernels.
//
// Arrays are the ghost-shaped staggered fields of the reference solver
// (C order, z fastest):
//   u (nx+1, ny+2, nz+2)   x-faces, y centres, z centres (+ghosts)
//   v (nx+2, ny+1, nz+2)   x centres, y-faces, z centres (+ghosts)
//   w (nx+2, ny+2, nz+1)   x centres, y centres, z faces (walls included)
//   p (nx+2, ny+2, nz+2)
// The interpolation "node buffer" of a component (semilag.py:387-394: one x-y
// period plus the whole z column) is read in place through NodeMap, so the
// reference's six node-buffer copies per advect call do not exist here.
//
// Floating point: every routine on the path arrival -> departure foot ->
// stencil index is written in the reference's operation order and this file
// is compiled with -ffp-contract=off, so the feet and the stencil indices are
// the same IEEE-754 results as the CPU oracle's.  Fused multiply-adds appear
// only where written explicitly (fma), i.e. in the high-order accumulation
// whose tolerance is a relative 1e-13, not bit-exactness.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace slc {

struct Grid {
    int nx, ny, nz;
    int order;        // field interpolation order P (4 | 6)
    int n_iters;      // iterated-midpoint fixed-point iterations
    int symmetric;    // 1: symmetric tanh map, 0: 'bottom'
    int top_neumann;  // top wall: 0 no-slip (odd ghost), 1 free-slip (even ghost)
    int xy_lagged;    // BDF2 predictor xy RHS: 0 2R^n - R^{n-1} ('extrap'), 1 R^n ('lagged')
    double dx, dy, Lz, gamma, tanh_g, z_lo, z_hi, nu;
    float rLz_f, rgamma_f, tanh_g_f;  // 1/Lz, 1/gamma, tanh(gamma) in float (z estimates)
    float split_f;  // slope of the centre-split estimate: -gamma/(2 nz) symmetric, gamma/(4 nz) bottom
    const double* z_c;      // nz+2 centres incl. ghost centres
    const double* z_f;      // nz+1 faces
    const double* dinv_c2;  // order-2 inverse Lagrange denominators, centre nodes
    const double* dinv_f2;  // order-2, face nodes
    const double* dinv_cP;  // order-P, centre nodes  ((nz+2)-P+1, P)
    const double* dinv_fP;  // order-P, face nodes    ((nz+1)-P+1, P)
    int traj_order;         // trajectory sampling order T: 2 trilinear, 4 tricubic
    const double* dinv_cT;  // order-T, centre nodes  ((nz+2)-T+1, T)
    const double* dinv_fT;  // order-T, face nodes    ((nz+1)-T+1, T)
};

// Component specs (semilag.py:282-286 node grids; solver.py:70-107 shapes).
template <int C>
struct Comp;

template <>
struct Comp<0> {  // u
    static constexpr double shift_x = 0.0, shift_y = -0.5;
    static constexpr bool centers = true;
    static constexpr int ox = 0, oy = 1;
    __device__ static long SY(const Grid& g) { return g.ny + 2; }
    __device__ static long SZ(const Grid& g) { return g.nz + 2; }
    __device__ static int NZ(const Grid& g) { return g.nz + 2; }
    __device__ static int NZI(const Grid& g) { return g.nz; }
};
template <>
struct Comp<1> {  // v
    static constexpr double shift_x = -0.5, shift_y = 0.0;
    static constexpr bool centers = true;
    static constexpr int ox = 1, oy = 0;
    __device__ static long SY(const Grid& g) { return g.ny + 1; }
    __device__ static long SZ(const Grid& g) { return g.nz + 2; }
    __device__ static int NZ(const Grid& g) { return g.nz + 2; }
    __device__ static int NZI(const Grid& g) { return g.nz; }
};
template <>
struct Comp<2> {  // w
    static constexpr double shift_x = -0.5, shift_y = -0.5;
    static constexpr bool centers = false;
    static constexpr int ox = 1, oy = 1;
    __device__ static long SY(const Grid& g) { return g.ny + 2; }
    __device__ static long SZ(const Grid& g) { return g.nz + 1; }
    __device__ static int NZ(const Grid& g) { return g.nz + 1; }
    __device__ static int NZI(const Grid& g) { return g.nz - 1; }
};

template <int C>
__device__ inline const double* znodes(const Grid& g) {
    return Comp<C>::centers ? g.z_c : g.z_f;
}

__device__ inline double clampd(double x, double lo, double hi) {
    double y = x < lo ? lo : x;  // torch.clamp: min(max(x, lo), hi)
    return y > hi ? hi : y;
}

__device__ inline int pymod(long a, int n) {
    long r = a % n;
    return (int)(r < 0 ? r + n : r);
}

// Python-style i mod n for 32-bit i when i lies more than one period outside
// [0, n) (feet displaced by more than the domain; non-finite feet).
__device__ inline int wrap_far(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}

// Python-style i mod n for 32-bit i.  Feet are CFL-bounded, so i is almost
// always within one period of [0, n): two selects (no divergent branch), and
// the division only for a lane outside [-n, 2n).
__device__ inline int wrap_idx(int i, int n) {
    int r = i < 0 ? i + n : i;
    r = r >= n ? r - n : r;
    if (__builtin_expect((unsigned)r >= (unsigned)n, 0)) r = wrap_far(i, n);
    return r;
}

// An integral-valued double (a floor) as an int.  The conversion instruction
// saturates out-of-range values (feet are CFL-bounded, so it never does).
__device__ inline int floor_int(double f) { return __double2int_rz(f); }

// Continuous face index kf (semilag.py:400-411).
__device__ inline double face_coord(const Grid& g, double z) {
    if (g.symmetric) {
        double arg = (2.0 * z / g.Lz - 1.0) * g.tanh_g;
        arg = clampd(arg, -1.0 + 1e-15, 1.0 - 1e-15);
        double xi = atanh(arg) / g.gamma;
        return (xi + 1.0) * (0.5 * (double)g.nz);
    } else {
        double arg = (1.0 - z / g.Lz) * g.tanh_g;
        arg = clampd(arg, -1.0 + 1e-15, 1.0 - 1e-15);
        double xi = 1.0 - atanh(arg) / g.gamma;
        return xi * (double)g.nz;
    }
}

// Face-interval index c = clamp(floor(kf), 0, nz-1) (semilag.py:415-416).
__device__ inline int face_cell(const Grid& g, double z) {
    double kf = face_coord(g, z);
    int c = floor_int(floor(kf));
    if (c < 0) c = 0;
    if (c > g.nz - 1) c = g.nz - 1;
    return c;
}

// The float estimate of the face interval of z that face_cell_fast starts
// from, clamped to [0, nz-1]; kfe receives the estimated face coordinate and
// a the map's tanh argument at z.
__device__ inline int face_estimate(const Grid& g, double z, float& kfe, float& a) {
    const int nz = g.nz;
    // float throughout, with approximate reciprocals: the estimate only has
    // to land within a few cells of the interval (the callers make it exact)
    const float zr = (float)z * g.rLz_f, tg = g.tanh_g_f, rg = g.rgamma_f;
    const float a0 = g.symmetric ? (2.0f * zr - 1.0f) * tg : (1.0f - zr) * tg;
    a = fminf(fmaxf(a0, -0.9999999f), 0.9999999f);
    const float at = 0.5f * __logf((1.0f + a) * __builtin_amdgcn_rcpf(1.0f - a));
    kfe = g.symmetric ? (at * rg + 1.0f) * (0.5f * (float)nz) : (1.0f - at * rg) * (float)nz;
    return (int)floorf(fminf(fmaxf(kfe, 0.0f), (float)(nz - 1)));
}

__device__ inline int face_estimate(const Grid& g, double z) {
    float kfe, a;
    return face_estimate(g, z, kfe, a);
}

// Predicted centre split of face interval c for a z with estimate (kfe, a):
// whether z >= z_c[c+1].  The centre is the faces' midpoint, which sits at
// 0.5 + (z''/z') h / 8 of the interval in the map's coordinate (h its face
// spacing) to first order; the slope is taken at z (a = tanh argument), so
// the prediction is off only within ~1e-4 of a cell of the centre.
__device__ inline bool split_predict(const Grid& g, int c, float kfe, float a) {
    return kfe - (float)c >= 0.5f + g.split_f * a;
}

// Whether face_cell_fast returns the estimate c as it is: z inside its faces
// lo = z_f[c], hi = z_f[c+1] (open at the clamped ends) and farther than
// 1e-10 of the cell height from both.
__device__ inline bool face_estimate_holds(const Grid& g, int c, double z, double lo,
                                           double hi) {
    const double band = 1e-10 * (hi - lo);
    const bool bracketed = (c == 0 || z >= lo) && (c == g.nz - 1 || z < hi);
    return bracketed && z - lo > band && hi - z > band;
}

// face_cell without the double-precision atanh.  A float32 estimate of the
// inverse map is refined against the face table to the interval
// z_f[c] <= z < z_f[c+1] (clamped to [0, nz-1]).  That interval is
// floor(kf(z)) except where kf(z) rounds across an integer, which can only
// happen for z within a few 1e-13 cell heights of a face (the analytic
// inverse of a face table entry is k to ~1e-13); within 1e-10 of a cell height
// of a face the analytic path itself runs, so the result always equals
// face_cell(g, z).  z must lie in [z_lo, z_hi].  split receives the centre
// node z_c[c+1] of the interval, loaded in the same round as its faces (for
// the estimated interval; again in the rare case the estimate moves).
__device__ inline int face_cell_fast(const Grid& g, double z, double& split) {
    const int nz = g.nz;
    int c = face_estimate(g, z);
    const double* zf = g.z_f;
    // both faces of the estimate at once (c + 1 <= nz is inside z_f); the
    // refinement loops only run when the estimate does not bracket z, which
    // is exactly when they would take a step
    double lo = zf[c], hi = zf[c + 1];
    split = g.z_c[c + 1];
    if (!((c == 0 || z >= lo) && (c == nz - 1 || z < hi))) {
        int guard = 0;
        while (c > 0 && z < zf[c] && guard < 8) { --c; ++guard; }
        while (c < nz - 1 && z >= zf[c + 1] && guard < 16) { ++c; ++guard; }
        lo = zf[c];
        hi = zf[c + 1];
        split = g.z_c[c + 1];
    }
    if (!face_estimate_holds(g, c, z, lo, hi)) {
        c = face_cell(g, z);
        split = g.z_c[c + 1];
    }
    return c;
}

__device__ inline int face_cell_fast(const Grid& g, double z) {
    double split;
    return face_cell_fast(g, z, split);
}

// Interval index m0 of component C's z nodes containing z
// (semilag.py:413-420), from the face interval c and its centre split.
template <int C>
__device__ inline int locate_split(int c, double split, double z) {
    if (!Comp<C>::centers) return c;
    return c + (z >= split ? 1 : 0);
}

// Interval index m0 of the component's z nodes containing z (semilag.py:413-420).
template <int C>
__device__ inline int locate_from_cell(const Grid& g, int c, double z) {
    if (!Comp<C>::centers) return c;
    return c + (z >= g.z_c[c + 1] ? 1 : 0);
}

// Uniform Lagrange weights on offsets l-(P/2-1) (semilag.py:426-439 with the
// inverse denominators of semilag.py:117-127), product taken in increasing l.
template <int P>
__device__ inline void uniform_weights(double t, double* w) {
    constexpr int h = P / 2 - 1;
    double d[P];
#pragma unroll
    for (int l = 0; l < P; ++l) d[l] = t - (double)(l - h);
#pragma unroll
    for (int m = 0; m < P; ++m) {
        double den = 1.0;
#pragma unroll
        for (int l = 0; l < P; ++l)
            if (l != m) den *= (double)((m - h) - (l - h));
        double num = 0.0;
        bool first = true;
#pragma unroll
        for (int l = 0; l < P; ++l) {
            if (l == m) continue;
            num = first ? d[l] : num * d[l];
            first = false;
        }
        w[m] = num * (1.0 / den);
    }
}

// Single uniform weight w[m] of uniform_weights<P> for a runtime m (same
// operation order, so the value is identical).
template <int P>
__device__ inline double uniform_weight_m(double t, int m) {
    constexpr int h = P / 2 - 1;
    double den = 1.0;
    for (int l = 0; l < P; ++l)
        if (l != m) den *= (double)((m - h) - (l - h));
    double num = 0.0;
    bool first = true;
    for (int l = 0; l < P; ++l) {
        if (l == m) continue;
        double d = t - (double)(l - h);
        num = first ? d : num * d;
        first = false;
    }
    return num * (1.0 / den);
}

// Nonuniform z weights against the actual nodes (semilag.py:441-454).
template <int P>
__device__ inline void z_weights(const double* zn, const double* dinv, double z, int k0,
                                 double* w) {
    double d[P];
#pragma unroll
    for (int l = 0; l < P; ++l) d[l] = z - zn[k0 + l];
#pragma unroll
    for (int m = 0; m < P; ++m) {
        double num = 0.0;
        bool first = true;
#pragma unroll
        for (int l = 0; l < P; ++l) {
            if (l == m) continue;
            num = first ? d[l] : num * d[l];
            first = false;
        }
        w[m] = num * dinv[k0 * P + m];
    }
}

// Linear address of node (ib, jb, k) of component C inside its ghost-shaped array.
// (32-bit: plans keep every field below 2^31 elements.  Signed on purpose:
// with no wrap-around the compiler proves base+k, base+k+1 adjacent and
// merges their loads into one 16-byte load.)
// PACKED: the address inside a packed node buffer (nx, ny, NZ) of the
// reference's _fill (semilag.py:387-394) instead.
template <int C, bool PACKED = false>
__device__ inline int node_addr(const Grid& g, int ib, int jb, int k) {
    if (PACKED) return (ib * g.ny + jb) * Comp<C>::NZ(g) + k;
    return ((ib + Comp<C>::ox) * (int)Comp<C>::SY(g) + (jb + Comp<C>::oy)) *
               (int)Comp<C>::SZ(g) + k;
}

// Arrival point of flat interior index n (z fastest) of component C
// (semilag.py:292-307).  Also returns the ghost-array indices (I, J, K).
template <int C>
__device__ inline void arrival(const Grid& g, long n, double& xa, double& ya, double& za, int& I,
                               int& J, int& K) {
    const unsigned nzi = (unsigned)Comp<C>::NZI(g);
    const unsigned nn = (unsigned)n;  // plans keep nx*ny*nz < 2^31
    int k = (int)(nn % nzi);
    unsigned r = nn / nzi;
    int j = (int)(r % (unsigned)g.ny);
    int i = (int)(r / (unsigned)g.ny);
    double ar = (double)(i + 1), aj = (double)(j + 1);
    if (C == 0) {
        xa = ar * g.dx;
        ya = (aj - 0.5) * g.dy;
        za = g.z_c[k + 1];
    } else if (C == 1) {
        xa = (ar - 0.5) * g.dx;
        ya = aj * g.dy;
        za = g.z_c[k + 1];
    } else {
        xa = (ar - 0.5) * g.dx;
        ya = (aj - 0.5) * g.dy;
        za = g.z_f[k + 1];
    }
    I = i + 1;
    J = j + 1;
    K = k + 1;
}

// Trajectory velocity accessor: either a stored field or U* = 2a - b formed
// on the fly (solver.py:866-868 forms the same single rounding).
struct TrajVel {
    static constexpr bool packed = false;
    const double* a[3];
    const double* b[3];  // null => U* = a
    __device__ inline double at(int c, long idx) const {
        return b[c] ? 2.0 * a[c][idx] - b[c][idx] : a[c][idx];
    }
};

// Trajectory velocity that is always a stored field (U* precomputed, or V^n).
struct TrajA {
    static constexpr bool packed = false;
    const double* a[3];
    __device__ inline double at(int c, long idx) const { return a[c][idx]; }
};

// Trajectory velocity read from the reference's packed node buffers
// (SLAdvector.mbuf, (nx, ny, NZ) per component, semilag.py:363-372, 604-606),
// the layout TritonSL.departure samples (semilag_triton.py:384-420).
struct TrajNodes {
    static constexpr bool packed = true;
    const double* a[3];
    __device__ inline double at(int c, long idx) const { return a[c][idx]; }
};

// Trilinear sample of trajectory component C at (x, y, zc) where the caller
// has already clamped z (semilag.py:512-528 with _build_iw_impl at order 2 and
// _gather_interp's sequential (i, j, k) sum).  `c` is the face cell of zc.
// TV: TrajVel or TrajA.
template <int C, typename TV = TrajVel>
__device__ inline double trilinear(const Grid& g, const TV& V, double xdx, double ydy,
                                   double zc, int cell) {
    double sx = xdx + Comp<C>::shift_x;
    double ix0 = floor(sx);
    double tx = sx - ix0;
    double sy = ydy + Comp<C>::shift_y;
    double iy0 = floor(sy);
    double ty = sy - iy0;
    const int fx = floor_int(ix0), fy = floor_int(iy0);
    int ixa = wrap_idx(fx, g.nx), ixb = wrap_idx(fx + 1, g.nx);
    int iya = wrap_idx(fy, g.ny), iyb = wrap_idx(fy + 1, g.ny);
    double wx[2], wy[2], wz[2];
    // order-2 uniform weights: (t - 1) * (-1), t * 1
    wx[0] = (tx - 1.0) * (-1.0);
    wx[1] = tx * 1.0;
    wy[0] = (ty - 1.0) * (-1.0);
    wy[1] = ty * 1.0;
    int m0 = locate_from_cell<C>(g, cell, zc);
    int k0 = m0;  // offset P/2-1 = 0 at order 2
    const int NZ = Comp<C>::NZ(g);
    if (k0 > NZ - 2) k0 = NZ - 2;
    if (k0 < 0) k0 = 0;
    const double* zn = znodes<C>(g);
    const double* dinv = Comp<C>::centers ? g.dinv_c2 : g.dinv_f2;
    double d0 = zc - zn[k0], d1 = zc - zn[k0 + 1];
    wz[0] = d1 * dinv[k0 * 2 + 0];
    wz[1] = d0 * dinv[k0 * 2 + 1];
    int ix[2] = {ixa, ixb}, iy[2] = {iya, iyb};
    double out = 0.0;
    bool first = true;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int base = node_addr<C, TV::packed>(g, ix[i], iy[j], k0);
            double wij = wx[i] * wy[j];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                double term = (wij * wz[k]) * V.at(C, base + k);
                out = first ? term : out + term;
                first = false;
            }
        }
    }
    return out;
}

// trilinear with the z node interval m0 given (the reference's locate
// result; k0 = m0 at order 2 -- no clamp is needed: m0 <= NZ - 2 for a
// clamped z).
// x/y part of the trilinear sample of component C: order-2 uniform weights
// ((t - 1) * (-1), t * 1) and the wrapped column indices.
template <int C>
__device__ inline void trilinear_xy(const Grid& g, double xdx, double ydy, double* wx,
                                    double* wy, int* ix, int* iy) {
    double sx = xdx + Comp<C>::shift_x;
    double ix0 = floor(sx);
    double tx = sx - ix0;
    double sy = ydy + Comp<C>::shift_y;
    double iy0 = floor(sy);
    double ty = sy - iy0;
    const int fx = floor_int(ix0), fy = floor_int(iy0);
    ix[0] = wrap_idx(fx, g.nx);
    ix[1] = wrap_idx(fx + 1, g.nx);
    iy[0] = wrap_idx(fy, g.ny);
    iy[1] = wrap_idx(fy + 1, g.ny);
    wx[0] = (tx - 1.0) * (-1.0);
    wx[1] = tx * 1.0;
    wy[0] = (ty - 1.0) * (-1.0);
    wy[1] = ty * 1.0;
}

template <int C, typename TV = TrajVel>
__device__ inline double trilinear_m(const Grid& g, const TV& V, double xdx, double ydy,
                                     double zc, int m0) {
    double wx[2], wy[2], wz[2];
    int ix[2], iy[2];
    trilinear_xy<C>(g, xdx, ydy, wx, wy, ix, iy);
    const int k0 = m0;
    const double* zn = znodes<C>(g);
    const double* dinv = Comp<C>::centers ? g.dinv_c2 : g.dinv_f2;
    double d0 = zc - zn[k0], d1 = zc - zn[k0 + 1];
    wz[0] = d1 * dinv[k0 * 2 + 0];
    wz[1] = d0 * dinv[k0 * 2 + 1];
    double out = 0.0;
    bool first = true;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int base = node_addr<C, TV::packed>(g, ix[i], iy[j], k0);
            double wij = wx[i] * wy[j];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                double term = (wij * wz[k]) * V.at(C, base + k);
                out = first ? term : out + term;
                first = false;
            }
        }
    }
    return out;
}

// Tricubic trajectory sample of component C (traj_interp_order = 4): the
// order-4 stencil of semilag.py:471-500 (no clamp counting; zc already
// clamped) and the sequential (i, j, k) sum of _gather_interp
// (semilag.py:44-59), the reference's operation order.
template <int C, typename TV = TrajVel>
__device__ inline double tricubic(const Grid& g, const TV& V, double xdx, double ydy,
                                  double zc, int cell) {
    constexpr int P = 4, h = 1;
    double sx = xdx + Comp<C>::shift_x;
    double ix0 = floor(sx);
    double tx = sx - ix0;
    double sy = ydy + Comp<C>::shift_y;
    double iy0 = floor(sy);
    double ty = sy - iy0;
    double wx[P], wy[P], wz[P];
    uniform_weights<P>(tx, wx);
    uniform_weights<P>(ty, wy);
    int k0 = locate_from_cell<C>(g, cell, zc) - h;
    const int NZ = Comp<C>::NZ(g);
    if (k0 < 0) k0 = 0;
    if (k0 > NZ - P) k0 = NZ - P;
    z_weights<P>(znodes<C>(g), Comp<C>::centers ? g.dinv_cT : g.dinv_fT, zc, k0, wz);
    const int fx = floor_int(ix0) - h, fy = floor_int(iy0) - h;
    double out = 0.0;
#pragma unroll 1
    for (int i = 0; i < P; ++i) {  // rolled: 16 loads in flight, not 64
        const int ix = wrap_idx(fx + i, g.nx);
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const int base = node_addr<C, TV::packed>(g, ix, wrap_idx(fy + j, g.ny), k0);
            double wij = wx[i] * wy[j];
#pragma unroll
            for (int k = 0; k < P; ++k) {
                double term = (wij * wz[k]) * V.at(C, base + k);
                out = (i == 0 && j == 0 && k == 0) ? term : out + term;
            }
        }
    }
    return out;
}

// Departure foot of one arrival point (semilag.py:534-552): K iterations of
// x_m <- x_a - (dt/2) U*(x_m), z_m clamped; returns x_d = 2 x_m - x_a with z
// unclamped.  The three trajectory samples at one point share x/dx, y/dy and
// the face cell of z (the reference recomputes them with identical results).
// T: trajectory sampling order (2 trilinear, 4 tricubic).
// TV: TrajVel (ghost-shaped fields) or TrajNodes (packed node buffers).
template <int T = 2, typename TV = TrajVel>
__device__ inline void foot(const Grid& g, const TV& V, double xa, double ya, double za,
                            double half_dt, double& xd, double& yd, double& zd) {
    double xm = xa, ym = ya, zm = za;
    for (int it = 0; it < g.n_iters; ++it) {
        double zc = clampd(zm, g.z_lo, g.z_hi);
        double xdx = xm / g.dx;
        double ydy = ym / g.dy;
        int cell = face_cell_fast(g, zc);
        double us, vs, ws;
        if constexpr (T == 4) {
            us = tricubic<0, TV>(g, V, xdx, ydy, zc, cell);
            vs = tricubic<1, TV>(g, V, xdx, ydy, zc, cell);
            ws = tricubic<2, TV>(g, V, xdx, ydy, zc, cell);
        } else {
            us = trilinear<0, TV>(g, V, xdx, ydy, zc, cell);
            vs = trilinear<1, TV>(g, V, xdx, ydy, zc, cell);
            ws = trilinear<2, TV>(g, V, xdx, ydy, zc, cell);
        }
        xm = xa - half_dt * us;
        ym = ya - half_dt * vs;
        zm = clampd(za - half_dt * ws, g.z_lo, g.z_hi);
    }
    xd = 2.0 * xm - xa;
    yd = 2.0 * ym - ya;
    zd = 2.0 * zm - za;
}

// Order-P stencil of a (clamped) foot (semilag.py:471-500).
template <int C, int P>
struct Stencil {
    int ix[P], iy[P];
    int k0;
    double wx[P], wy[P], wz[P];
};

template <int C, int P>
__device__ inline void build_stencil(const Grid& g, double x, double y, double z,
                                     Stencil<C, P>& s) {
    constexpr int h = P / 2 - 1;
    double sx = x / g.dx + Comp<C>::shift_x;
    double ix0 = floor(sx);
    double tx = sx - ix0;
    double sy = y / g.dy + Comp<C>::shift_y;
    double iy0 = floor(sy);
    double ty = sy - iy0;
    int x0 = wrap_idx(floor_int(ix0) - h, g.nx), y0 = wrap_idx(floor_int(iy0) - h, g.ny);
#pragma unroll
    for (int l = 0; l < P; ++l) {
        int xi = x0 + l;
        while (xi >= g.nx) xi -= g.nx;  // at most one pass unless nx < P
        s.ix[l] = xi;
        int yi = y0 + l;
        while (yi >= g.ny) yi -= g.ny;
        s.iy[l] = yi;
    }
    uniform_weights<P>(tx, s.wx);
    uniform_weights<P>(ty, s.wy);
    int m0 = locate_from_cell<C>(g, face_cell_fast(g, z), z);
    int k0 = m0 - h;
    const int NZ = Comp<C>::NZ(g);
    if (k0 < 0) k0 = 0;
    if (k0 > NZ - P) k0 = NZ - P;
    s.k0 = k0;
    z_weights<P>(znodes<C>(g), Comp<C>::centers ? g.dinv_cP : g.dinv_fP, z, k0, s.wz);
}

// Staged tensor-product contraction (z, then y, then x) with FMAs.
template <int C, int P>
__device__ inline double gather(const Grid& g, const double* __restrict__ F,
                                const Stencil<C, P>& s) {
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < P; ++i) {
        double ai = 0.0;
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const double* col = F + node_addr<C>(g, s.ix[i], s.iy[j], s.k0);
            double sj = 0.0;
#pragma unroll
            for (int k = 0; k < P; ++k) sj = fma(s.wz[k], col[k], sj);
            ai = fma(s.wy[j], sj, ai);
        }
        acc = fma(s.wx[i], ai, acc);
    }
    return acc;
}

// Unwrapped integer stencil base of a (clamped) foot: x/y node index of stencil
// offset 0 before the periodic wrap, and the clamped z base k0.  Same
// arithmetic as build_stencil.
template <int C, int P>
__device__ inline void stencil_base(const Grid& g, double x, double y, double z, int& bx,
                                    int& by, int& k0, double& tx, double& ty) {
    constexpr int h = P / 2 - 1;
    double sx = x / g.dx + Comp<C>::shift_x;
    double ix0 = floor(sx);
    tx = sx - ix0;
    double sy = y / g.dy + Comp<C>::shift_y;
    double iy0 = floor(sy);
    ty = sy - iy0;
    bx = floor_int(ix0) - h;
    by = floor_int(iy0) - h;
    double split;
    const int c = face_cell_fast(g, z, split);
    int m0 = locate_split<C>(c, split, z);
    int kb = m0 - h;
    const int NZ = Comp<C>::NZ(g);
    if (kb < 0) kb = 0;
    if (kb > NZ - P) kb = NZ - P;
    k0 = kb;
}

// Explicit xy-diffusion nu*(d2x + d2y) at ghost index (I, J, K)
// (operators.py:293-367; periodic extension for u in x and v in y).
// rdx2, rdy2: 1/dx^2, 1/dy^2 (a multiply in place of the reference's
// division: this term carries the predictor's 1e-13 tolerance, not the
// bit-exact feet)
template <int C>
__device__ inline double rhs_xy(const Grid& g, const double* __restrict__ F, int I, int J, int K,
                                double rdx2, double rdy2) {
    const long SY = Comp<C>::SY(g), SZ = Comp<C>::SZ(g);
    int Ip = I + 1, Jp = J + 1;
    if (C == 0 && Ip == g.nx + 1) Ip = 1;
    if (C == 1 && Jp == g.ny + 1) Jp = 1;
    auto at = [&](int i, int j) { return F[((long)i * SY + j) * SZ + K]; };
    double c = at(I, J);
    double d2x = ((at(Ip, J) - 2.0 * c) + at(I - 1, J)) * rdx2;
    double d2y = ((at(I, Jp) - 2.0 * c) + at(I, J - 1)) * rdy2;
    return g.nu * (d2x + d2y);
}

// blockIdx -> tile remap that keeps consecutive tiles on one XCD (blocks are
// dispatched round-robin over the 8 XCDs; each XCD has its own L2).
__device__ inline long xcd_block(long nblocks) {
    long b = blockIdx.x;
    const long nx = 8;
    long per = nblocks / nx;
    long rem = nblocks % nx;
    long x = b % nx, q = b / nx;
    // blocks of XCD x occupy logical range [start_x, start_x + count_x)
    long start = x * per + (x < rem ? x : rem);
    long count = per + (x < rem ? 1 : 0);
    return q < count ? start + q : b;
}

}  // namespace slc
