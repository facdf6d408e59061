/*
 * sl_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY) for the semi-Lagrangian
 * machinery of slChannel: departure feet and tensor-product Lagrange gathers.
 *
 * This file restates, point by point and in the same floating-point operation
 * order, the reference's eager fp64 path in slchannel/semilag.py.  It is the
 * checker for the HIP kernels, never the product: only tests, the smoke check
 * and the CPU-baseline leg of the benchmark load it.
 *
 * Build (see oracle/Makefile): gcc -O2 -ffp-contract=off -fno-fast-math, so
 * every + - * / is one IEEE-754 binary64 operation exactly as torch's eager
 * CPU kernels perform it.  The only library call on the index path is atanh
 * (glibc here, SLEEF-like in torch); the two can differ by one ulp, which can
 * only move floor(kf) when a point sits exactly on a z face (a "tie").
 *
 * Fields are the ghost-shaped staggered arrays (C order, z fastest):
 *   u (nx+1, ny+2, nz+2), v (nx+2, ny+1, nz+2), w (nx+2, ny+2, nz+1).
 * The reference copies one x-y period of each into a "node buffer"
 * (semilag.py:387-394); reading the ghost-shaped array through the same index
 * map is the identical value, so no copy is made here.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

typedef struct {
    int nx, ny, nz;
    int order;        /* field interpolation order P (4|6) */
    int n_iters;      /* midpoint iterations */
    int symmetric;    /* 1 = 'symmetric' tanh map, 0 = 'bottom' */
    double dx, dy, Lz, gamma, tanh_g, z_lo, z_hi;
    const double* z_c;      /* nz+2 cell centres incl. ghosts */
    const double* z_f;      /* nz+1 faces */
    const double* dinv_c2;  /* (nz+2-2+1, 2) inverse denominators, centre nodes */
    const double* dinv_f2;  /* (nz+1-2+1, 2) face nodes */
    const double* dinv_cP;  /* (nz+2-P+1, P) */
    const double* dinv_fP;  /* (nz+1-P+1, P) */
    int traj_order;         /* trajectory sampling order T (2 = trilinear | 4 = tricubic) */
    const double* dinv_cT;  /* (nz+2-T+1, T) */
    const double* dinv_fT;  /* (nz+1-T+1, T) */
} orc_grid;

/* Node-grid spec of one component (semilag.py:282-286). */
typedef struct {
    double shift_x, shift_y;
    int NX, NY, NZ;
    int centers;
    const double* znodes;
    /* ghost-shaped addressing of node (ib, jb, k): ((ib+ox)*SY + jb+oy)*SZ + k */
    int ox, oy;
    long SY, SZ;
} orc_spec;

static orc_spec make_spec(const orc_grid* g, int comp) {
    orc_spec s;
    s.NX = g->nx;
    s.NY = g->ny;
    if (comp == 0) {        /* u: x-faces, y centres, z centres */
        s.shift_x = 0.0; s.shift_y = -0.5; s.centers = 1; s.znodes = g->z_c;
        s.NZ = g->nz + 2; s.ox = 0; s.oy = 1; s.SY = g->ny + 2; s.SZ = g->nz + 2;
    } else if (comp == 1) { /* v */
        s.shift_x = -0.5; s.shift_y = 0.0; s.centers = 1; s.znodes = g->z_c;
        s.NZ = g->nz + 2; s.ox = 1; s.oy = 0; s.SY = g->ny + 1; s.SZ = g->nz + 2;
    } else {                /* w: z faces, walls included */
        s.shift_x = -0.5; s.shift_y = -0.5; s.centers = 0; s.znodes = g->z_f;
        s.NZ = g->nz + 1; s.ox = 1; s.oy = 1; s.SY = g->ny + 2; s.SZ = g->nz + 1;
    }
    return s;
}

static inline double clampd(double x, double lo, double hi) {
    /* torch.clamp(x, lo, hi) == min(max(x, lo), hi) */
    double y = x < lo ? lo : x;
    return y > hi ? hi : y;
}

static inline long pymod(long a, long n) {
    long r = a % n;
    return r < 0 ? r + n : r;
}

/* semilag.py:400-411 (_face_coord) */
double orc_face_coord(const orc_grid* g, double z) {
    if (g->symmetric) {
        double arg = (2.0 * z / g->Lz - 1.0) * g->tanh_g;
        arg = clampd(arg, -1.0 + 1e-15, 1.0 - 1e-15);
        double xi = atanh(arg) / g->gamma;
        return (xi + 1.0) * (0.5 * (double)g->nz);
    } else {
        double arg = (1.0 - z / g->Lz) * g->tanh_g;
        arg = clampd(arg, -1.0 + 1e-15, 1.0 - 1e-15);
        double xi = 1.0 - atanh(arg) / g->gamma;
        return xi * (double)g->nz;
    }
}

/* semilag.py:413-420 (_locate_z) */
static long locate_z(const orc_grid* g, const orc_spec* s, double z) {
    double kf = orc_face_coord(g, z);
    long c = (long)floor(kf);
    if (c < 0) c = 0;
    if (c > g->nz - 1) c = g->nz - 1;
    if (!s->centers) return c;
    return c + (z >= s->znodes[c + 1] ? 1 : 0);
}

/* _locate_z for n points of component comp (z clamped into [z_lo, z_hi]
 * first, as every caller in the reference does). */
void orc_locate_z(const orc_grid* g, int comp, long n, const double* z, int32_t* m0) {
    orc_spec s = make_spec(g, comp);
    for (long i = 0; i < n; ++i) m0[i] = (int32_t)locate_z(g, &s, clampd(z[i], g->z_lo, g->z_hi));
}

/* semilag.py:426-439 (_uniform_weights) with semilag.py:117-127 baked in */
static void uniform_weights(double t, int P, double* w) {
    double d[8], inv[8];
    int h = P / 2 - 1;
    for (int m = 0; m < P; ++m) {
        double den = 1.0;
        for (int l = 0; l < P; ++l)
            if (l != m) den *= (double)((m - h) - (l - h));
        inv[m] = 1.0 / den;
    }
    for (int l = 0; l < P; ++l) d[l] = t - (double)(l - h);
    for (int m = 0; m < P; ++m) {
        double num = 0.0;
        int first = 1;
        for (int l = 0; l < P; ++l) {
            if (l == m) continue;
            if (first) { num = d[l]; first = 0; }
            else num = num * d[l];
        }
        w[m] = num * inv[m];
    }
}

/* semilag.py:441-454 (_z_weights) */
static void z_weights(const orc_spec* s, const double* dinv, double z, long k0, int P,
                      double* w) {
    double d[8];
    for (int l = 0; l < P; ++l) d[l] = z - s->znodes[k0 + l];
    for (int m = 0; m < P; ++m) {
        double num = 0.0;
        int first = 1;
        for (int l = 0; l < P; ++l) {
            if (l == m) continue;
            if (first) { num = d[l]; first = 0; }
            else num = num * d[l];
        }
        w[m] = num * dinv[k0 * P + m];
    }
}

typedef struct {
    long ix[8], iy[8];  /* node indices after the periodic wrap */
    long k0;
    double wx[8], wy[8], wz[8];
} stencil;

/* semilag.py:471-500 (_build_iw_impl), without the clamp (done by callers) */
static void build_stencil(const orc_grid* g, const orc_spec* s, const double* dinv, double x,
                          double y, double z, int P, stencil* st) {
    int h = P / 2 - 1;
    double sx = x / g->dx + s->shift_x;
    double ix0 = floor(sx);
    double tx = sx - ix0;
    for (int l = 0; l < P; ++l) st->ix[l] = pymod((long)ix0 + (l - h), s->NX);
    uniform_weights(tx, P, st->wx);

    double sy = y / g->dy + s->shift_y;
    double iy0 = floor(sy);
    double ty = sy - iy0;
    for (int l = 0; l < P; ++l) st->iy[l] = pymod((long)iy0 + (l - h), s->NY);
    uniform_weights(ty, P, st->wy);

    long m0 = locate_z(g, s, z);
    long k0 = m0 - h;
    if (k0 < 0) k0 = 0;
    if (k0 > s->NZ - P) k0 = s->NZ - P;
    st->k0 = k0;
    z_weights(s, dinv, z, k0, P, st->wz);
}

/* semilag.py:44-59 (_gather_interp): sequential (i, j, k) accumulation */
static double gather(const orc_spec* s, const double* F, const stencil* st, int P) {
    double out = 0.0;
    int first = 1;
    for (int i = 0; i < P; ++i) {
        for (int j = 0; j < P; ++j) {
            long base = ((st->ix[i] + s->ox) * s->SY + (st->iy[j] + s->oy)) * s->SZ + st->k0;
            double wij = st->wx[i] * st->wy[j];
            for (int k = 0; k < P; ++k) {
                double term = (wij * st->wz[k]) * F[base + k];
                if (first) { out = term; first = 0; }
                else out = out + term;
            }
        }
    }
    return out;
}

/* semilag.py:512-528 (_sample): trajectory velocity at order traj_order
 * (2 = trilinear, 4 = tricubic): the order-T stencil without clamp counting
 * and the sequential gather */
static double sample(const orc_grid* g, int comp, const double* F, double x, double y,
                     double z) {
    orc_spec s = make_spec(g, comp);
    const int T = g->traj_order;
    const double* dinv = s.centers ? g->dinv_cT : g->dinv_fT;
    stencil st;
    build_stencil(g, &s, dinv, x, y, clampd(z, g->z_lo, g->z_hi), T, &st);
    return gather(&s, F, &st, T);
}

/* Arrival coordinates (semilag.py:292-307) of the flat interior index n. */
static void arrival(const orc_grid* g, int comp, long n, double* xa, double* ya, double* za) {
    int nzi = comp == 2 ? g->nz - 1 : g->nz;
    long k = n % nzi;
    long j = (n / nzi) % g->ny;
    long i = n / ((long)nzi * g->ny);
    double ar = (double)(i + 1), aj = (double)(j + 1);
    if (comp == 0) {
        *xa = ar * g->dx; *ya = (aj - 0.5) * g->dy; *za = g->z_c[k + 1];
    } else if (comp == 1) {
        *xa = (ar - 0.5) * g->dx; *ya = aj * g->dy; *za = g->z_c[k + 1];
    } else {
        *xa = (ar - 0.5) * g->dx; *ya = (aj - 0.5) * g->dy; *za = g->z_f[k + 1];
    }
}

long orc_n_arrival(const orc_grid* g, int comp) {
    return (long)g->nx * g->ny * (comp == 2 ? g->nz - 1 : g->nz);
}

/* semilag.py:534-552 (departure_coords); z of the foot is NOT clamped here */
void orc_departure(const orc_grid* g, int comp, const double* um, const double* vm,
                   const double* wm, double dt, double* xd, double* yd, double* zd) {
    long N = orc_n_arrival(g, comp);
    double half_dt = 0.5 * dt;
#pragma omp parallel for schedule(static)
    for (long n = 0; n < N; ++n) {
        double xa, ya, za;
        arrival(g, comp, n, &xa, &ya, &za);
        double xm = xa, ym = ya, zm = za;
        for (int it = 0; it < g->n_iters; ++it) {
            double us = sample(g, 0, um, xm, ym, zm);
            double vs = sample(g, 1, vm, xm, ym, zm);
            double ws = sample(g, 2, wm, xm, ym, zm);
            xm = xa - half_dt * us;
            ym = ya - half_dt * vs;
            zm = clampd(za - half_dt * ws, g->z_lo, g->z_hi);
        }
        xd[n] = 2.0 * xm - xa;
        yd[n] = 2.0 * ym - ya;
        zd[n] = 2.0 * zm - za;
    }
}

/* semilag.py:471-474 + 496-500 + 44-59: clamp-and-count the foot z, build the
 * order-P stencil and gather F.  Optionally reports the stencil base indices
 * (ix of offset -(P/2-1), iy likewise, k0).  Returns the clamp count. */
long orc_gather(const orc_grid* g, int comp, const double* F, const double* xd,
                const double* yd, const double* zd, double* out, int32_t* ix_out,
                int32_t* iy_out, int32_t* k0_out) {
    long N = orc_n_arrival(g, comp);
    int P = g->order;
    long ncl = 0;
#pragma omp parallel for schedule(static) reduction(+ : ncl)
    for (long n = 0; n < N; ++n) {
        orc_spec s = make_spec(g, comp);
        const double* dinv = s.centers ? g->dinv_cP : g->dinv_fP;
        double z = zd[n];
        if (z < g->z_lo || z > g->z_hi) ncl += 1;
        z = clampd(z, g->z_lo, g->z_hi);
        stencil st;
        build_stencil(g, &s, dinv, xd[n], yd[n], z, P, &st);
        out[n] = gather(&s, F, &st, P);
        if (ix_out) ix_out[n] = (int32_t)st.ix[0];
        if (iy_out) iy_out[n] = (int32_t)st.iy[0];
        if (k0_out) k0_out[n] = (int32_t)st.k0;
    }
    return ncl;
}

/* Trajectory sample (order traj_order) at arbitrary points
 * (semilag.py:512-528), for tests of the sampling alone. */
void orc_sample(const orc_grid* g, int comp, const double* F, long n, const double* x,
                const double* y, const double* z, double* out) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) out[i] = sample(g, comp, F, x[i], y[i], z[i]);
}

/* Order-P interpolation at arbitrary points (semilag.py:460-469 + 502-510),
 * z assumed inside [z_lo, z_hi]; used by the interpolation-convergence tests. */
void orc_interp(const orc_grid* g, int comp, const double* F, long n, const double* x,
                const double* y, const double* z, double* out) {
    int P = g->order;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        orc_spec s = make_spec(g, comp);
        const double* dinv = s.centers ? g->dinv_cP : g->dinv_fP;
        stencil st;
        build_stencil(g, &s, dinv, x[i], y[i], z[i], P, &st);
        out[i] = gather(&s, F, &st, P);
    }
}
