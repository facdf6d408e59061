/*
 * slc.h -- C ABI of libslc_hip.so, the MI355X (gfx950) performance layer for
 * slChannel's semi-Lagrangian BDF2 step and the pressure projection around it.
 *
 * Conventions
 *   - Every entry point returns an int status: 0 (SLC_OK) or a negative
 *     SLC_E* code; nothing throws across the ABI.  slc_last_error() gives a
 *     human-readable message for the calling thread.
 *   - Array arguments are DEVICE pointers to float64 data owned by the caller
 *     (borrowed, never freed).  Fields are the reference's ghost-shaped
 *     staggered arrays, C order with z fastest:
 *        u (nx+1, ny+2, nz+2)   v (nx+2, ny+1, nz+2)
 *        w (nx+2, ny+2, nz+1)   p (nx+2, ny+2, nz+2)
 *   - `stream` is a hipStream_t (NULL = the legacy default stream).  All work
 *     is enqueued on it; no entry point synchronises the host, so results
 *     (including the clamp counter and the bulk scalars) stay on the device.
 *   - Components are numbered 0 = u, 1 = v, 2 = w.
 *
 * Each function names the reference interface (gmcavallazzi/SLchannel,
 * file:line) it replaces.
 */
#ifndef SLC_H
#define SLC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SLC_OK 0
#define SLC_EINVAL (-1)   /* bad argument / unsupported option */
#define SLC_EHIP (-2)     /* HIP runtime error */
#define SLC_EFFT (-3)     /* rocFFT error */
#define SLC_ENOMEM (-4)   /* device allocation failed */

typedef struct slc_plan slc_plan;

/* Grid and scheme description.  Mirrors the arguments of
 * SLAdvector.__init__ (slchannel/semilag.py:133-153) plus the solver scalars
 * the step needs (slchannel/solver.py:195-255, 311-321). */
typedef struct {
    int nx, ny, nz;        /* interior cells; x, y periodic, z wall-bounded */
    double Lx, Ly, Lz;     /* dx = Lx/nx, dy = Ly/ny (solver.py:315-316) */
    double gamma;          /* tanh stretching factor */
    double tanh_g;         /* tanh(gamma) as the host computed it (semilag.py:240-242) */
    int stretching;        /* 0 = 'symmetric', 1 = 'bottom' (semilag.py:210-213) */
    int order;             /* 4 | 6 (semilag.py:204-205) */
    int traj_order;        /* trajectory sampling: 2 trilinear (the LDS-tiled step
                              kernels), 4 tricubic (direct-load kernels);
                              semilag.py:179-209 */
    int n_traj_iters;      /* midpoint fixed-point iterations (semilag.py:543) */
    int top_wall;          /* 0 = 'dirichlet' (no-slip), 1 = 'neumann' (free-slip) */
    double nu;             /* 1/Re (solver.py:209) */
    const double* z_f;     /* HOST pointer, nz+1 faces   (utils.py:21-61) */
    const double* z_c;     /* HOST pointer, nz+2 centres (utils.py:21-61) */
} slc_grid_desc;

/* Library version (major*10000 + minor*100 + patch). */
int slc_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char* slc_last_error(void);

/* Build a plan on HIP device `device`: uploads the z tables, derives the
 * Lagrange denominator tables (semilag.py:102-127), the Poisson tables
 * (projection.py:24-131, kept as 1-D factors, not (nx, ny/2+1, nz) arrays),
 * creates the rocFFT plans and the work buffers.
 * Replaces SLAdvector.__init__ (semilag.py:133-381), TritonSL.__init__
 * (semilag_triton.py:352-382) and initialize_fft_solver (projection.py:24-131). */
int slc_plan_create(const slc_grid_desc* desc, int device, slc_plan** out);

/* Free a plan and its device buffers. */
int slc_plan_destroy(slc_plan* plan);

/* Device bytes held by the plan (tables + work buffers). */
int slc_plan_workspace_bytes(const slc_plan* plan, int64_t* out);

/* Ghost refresh of the three velocity fields (any may be NULL).
 * Replaces apply_bc_all (slchannel/solver.py:70-107). */
int slc_apply_bc(slc_plan* plan, double* u, double* v, double* w, void* stream);

/* Departure feet of every arrival point of component `comp` by the iterated
 * midpoint rule, sampling the trajectory velocity (um, vm, wm) at the plan's
 * traj_order (2 trilinear, 4 tricubic).  layout 0: um, vm, wm are the
 * ghost-shaped fields; layout 1: the packed node buffers (nx, ny, NZ) that
 * SLAdvector._fill writes into SLAdvector.mbuf (semilag.py:387-394, 604-606;
 * NZ = nz+2 for u, v and nz+1 for w).  Outputs are flat arrays of nx*ny*nzi
 * doubles (nzi = nz for u, v and nz-1 for w, z fastest); x, y unwrapped.
 * nclamp NULL: z is NOT clamped (SLAdvector.departure_coords,
 * semilag.py:534-552).  nclamp non-NULL (a device int64): z is clamped into
 * [z_lo, z_hi] and the number of clamped feet is added to *nclamp
 * (TritonSL.departure, semilag_triton.py:384-420, 262-269); a later
 * slc_gather then needs no counter.
 * Replaces SLAdvector.departure_coords / TritonSL.departure. */
int slc_departure(slc_plan* plan, int comp, const double* um, const double* vm,
                  const double* wm, int layout, double dt, double* xd, double* yd, double* zd,
                  int64_t* nclamp, void* stream);

/* Clamp the feet into [z_lo, z_hi] (adding the number clamped to *nclamp, a
 * device int64, if non-NULL), build the order-P stencils and interpolate
 * `field`.  layout 0: `field` is the component's ghost-shaped array;
 * layout 1: a packed node buffer (nx, ny, NZ) as built by SLAdvector._fill
 * (NZ = nz+2 for u, v; nz+1 for w).  out: flat nx*ny*nzi.
 * Replaces _build_iw_impl + _apply_iw (semilag.py:471-510) /
 * TritonSL.gather (semilag_triton.py:422-455). */
int slc_gather(slc_plan* plan, int comp, const double* field, int layout, const double* xd,
               const double* yd, const double* zd, double* out, int64_t* nclamp,
               void* stream);

/* Stencil base indices of given feet after the clamp: ix/iy = node index of
 * the first stencil point (offset -(P/2-1)) after the periodic wrap, k0 = the
 * clamped z base (int32 arrays of nx*ny*nzi).  Adds the clamp count to
 * *nclamp if non-NULL.  Replaces the index half of
 * SLAdvector.compute_departure -> IndexWeights (semilag.py:471-500, 573-578). */
int slc_stencil(slc_plan* plan, int comp, const double* xd, const double* yd, const double* zd,
                int32_t* ix, int32_t* iy, int32_t* k0, int64_t* nclamp, void* stream);

/* z-locate of n points (device doubles, clamped into [z_lo, z_hi] first):
 * node-interval index m0 of component comp's z nodes, by the analytic inverse
 * tanh map (m0_exact) and by the table-refined fast path the kernels use
 * (m0_fast).  Replaces SLAdvector._locate_z (semilag.py:413-420). */
int slc_locate_z(slc_plan* plan, int comp, const double* z, int64_t n, int32_t* m0_exact,
                 int32_t* m0_fast, void* stream);

/* Full semi-Lagrangian advection of (u, v, w) along (um, vm, wm) over dt,
 * writing the interiors of the ghost-shaped outputs (ghosts and the w wall
 * faces untouched) and adding the clamp count to *nclamp (device int64,
 * may be NULL).  Replaces SLAdvector.advect (semilag.py:584-671). */
int slc_advect(slc_plan* plan, const double* u, const double* v, const double* w,
               const double* um, const double* vm, const double* wm, double dt,
               double* ustar, double* vstar, double* wstar, int64_t* nclamp, void* stream);

/* BDF1 bootstrap predictor: ustar = advect(V^n; U* = V^n, dt) + dt * nu *
 * L_xy V^n on the interiors.  Replaces solver.py:737-764. */
int slc_bdf1_predict(slc_plan* plan, const double* u, const double* v, const double* w,
                     double dt, double* ustar, double* vstar, double* wstar, int64_t* nclamp,
                     void* stream);

/* BDF2 predictor with the frozen U* = 2V^n - V^{n-1}: two independent feet
 * (dt, 2dt), V^n interpolated at the near foot and V^{n-1} at the far foot,
 * u_hat = 4/3 near - 1/3 far + (2dt/3)(2 R^n - R^{n-1}) with R = nu L_xy V.
 * Clamp counts of both feet are added to *nclamp.
 * Replaces solver.py:846-903 ('noninc' pressure, 'extrap' xy RHS). */
int slc_bdf2_predict(slc_plan* plan, const double* u, const double* v, const double* w,
                     const double* unm1, const double* vnm1, const double* wnm1, double dt,
                     double* ustar, double* vstar, double* wstar, int64_t* nclamp,
                     void* stream);

/* In-place theta-method implicit z-diffusion of one component's interior
 * (ghosts untouched).  Replaces solve_implicit_diffusion_{u,v,w}
 * (slchannel/operators.py:382-630). */
int slc_zsolve(slc_plan* plan, int comp, double* field, double dt, double theta, void* stream);

/* Divergence of the velocity into a (nx, ny, nz) z-contiguous array.
 * Replaces compute_divergence (slchannel/utils.py:166-180). */
int slc_divergence(slc_plan* plan, const double* u, const double* v, const double* w,
                   double* div, void* stream);

/* Solve the pressure Poisson equation for rhs = div / rhs_scale (div a
 * (nx, ny, nz) z-contiguous array) into the ghost-shaped p.
 * Replaces solve_poisson_fft (slchannel/projection.py:133-197). */
int slc_poisson(slc_plan* plan, const double* div, double rhs_scale, double* p, void* stream);

/* u -= dt grad p on the interiors.  Replaces project_velocity
 * (slchannel/projection.py:250-272). */
int slc_project(slc_plan* plan, double* u, double* v, double* w, const double* p, double dt,
                void* stream);

/* out = a x + b y over the plan's pressure array ((nx+2)(ny+2)(nz+2)
 * doubles), each product rounded, then the sum: the incremental-pressure
 * bookkeeping of the 'inc' BDF2 option, p_ext = 2 p^n - p^{n-1} (a = 2,
 * b = -1; solver.py:913-917) and p^{n+1} = p_ext + phi (a = b = 1;
 * solver.py:999-1011).  out may alias x or y. */
int slc_pressure_combine(slc_plan* plan, double* out, double a, const double* x, double b,
                         const double* y, void* stream);

/* The projection block of one step, fused: div(u)/rhs_scale -> rocFFT ->
 * per-mode Thomas -> inverse rocFFT -> u -= dt_proj grad p -> ghosts of
 * u, v, w and p.  p may be NULL (pressure not materialised).
 * Replaces solver.py:983-998 + 1015 (and 797-801 + 809 in the bootstrap). */
int slc_project_step(slc_plan* plan, double* u, double* v, double* w, double rhs_scale,
                     double dt_proj, double* p, void* stream);

/* Bulk velocity sum(u dx dy dz_f)/V of u's interior into out[3] (out: 4
 * device doubles of scratch).  Replaces compute_bulk_velocity
 * (slchannel/utils.py:147-163). */
int slc_bulk_velocity(slc_plan* plan, const double* u, double* out, void* stream);

/* Eulerian IMEX explicit terms, compute_momentum_rhs_fused_imex
 * (operators.py:639-835; eulerian_triton.py:31-193): divergence-form
 * advection and the xy Laplacian, rhs = nu (d2x + d2y) - div(f V) on each
 * component's interior, 0 on every other entry of the ghost-shaped outputs
 * ru, rv, rw (same shapes as u, v, w).  Each element in the reference's
 * operation order. */
int slc_imex_rhs(slc_plan* plan, const double* u, const double* v, const double* w, double* ru,
                 double* rv, double* rw, void* stream);

/* AB2 explicit update of step_imex (solver.py:1074-1081) over the whole
 * arrays: f += dt (1.5 r - 0.5 r_prev), or f += dt r when the three r_prev
 * are null (the first step). */
int slc_ab2_update(slc_plan* plan, double* u, double* v, double* w, const double* ru,
                   const double* rv, const double* rw, const double* ru_prev,
                   const double* rv_prev, const double* rw_prev, double dt, void* stream);

/* Run-loop diagnostics of the current fields, written to out (5 device
 * doubles, no host synchronisation): [0] max|div| (compute_divergence,
 * utils.py:166-180), [1] u_tau (compute_u_tau, utils.py:116-146), [2] the
 * advective CFL dti * dt and [3] dti (compute_cfl_fused,
 * operators.py:839-940), [4] the number of NaN/Inf entries of u, v, w and p
 * (the run loop's guard, solver.py:1436-1479; p may be null).  Replaces the
 * per-n_out diagnostics of the run loop (solver.py:1355-1392); the maxima
 * propagate NaN. */
int slc_diagnostics(slc_plan* plan, const double* u, const double* v, const double* w,
                    const double* p, double dt, double* out, void* stream);

/* Seeded solenoidal initial velocity on the device (no counterpart in the
 * reference, whose initialisers are the parabolic / vortex / uniform fields of
 * initflow.py:123-223; the seeded perturbation of SLChannelFlow.initialize is
 * this library's own).  Writes the interiors of u, v, w (ghost-shaped; ghosts
 * untouched): the curl of A_c = sum_b Re[a^c_b e^{i(kx x + ky y)}] G_l(z)
 * over W x W wavenumber bins (W odd, bins (m, n) with |m|, |n| <= (W-1)/2)
 * and l = 1..L, scaled to the given u'rms over the three interiors, plus
 * the base profile on u.  Device inputs:
 *   ab     five (W*W, L) complex blocks (re, im interleaved): the
 *          coefficients of G_l and dG_l/dz per bin -- alpha_u, beta_u,
 *          alpha_v, beta_v, alpha_w;
 *   zt     G_l at the interior centres (L, nz), dG_l/dz there (L, nz), G_l at
 *          the interior faces (L, nz-1), the base profile of u (nz);
 *   phase  per component (u, v, w) the x factors e^{i kx_m x_i} (W, nx), then
 *          the y factors e^{i ky_n y_j} (W, ny), at its staggered positions;
 *   work   scratch of at least slc_init_solenoidal_work(nx, ny, nz, W) bytes.
 * Four launches on the stream, no host synchronisation. */
int slc_init_solenoidal(slc_plan* plan, int W, int L, const double* ab, const double* zt,
                        const double* phase, double urms, double* work, int64_t work_bytes,
                        double* u, double* v, double* w, void* stream);
int64_t slc_init_solenoidal_work(int nx, int ny, int nz, int W);

/* Trilinear interpolation of a source field src (sx x sy x sz doubles, z
 * contiguous) onto nxq x nyq x nzq tensor-product query nodes, given per axis
 * as the lower source index and the weight of the upper node (ix[a] with
 * ix[a] + 1 < sx, tx[a] in [0, 1]; likewise y and z; device arrays):
 * out[a][b][c] = sum_{p,q,r} ((wx_p wy_q) wz_r) src[ix+p][iy+q][iz+r],
 * wx_0 = 1 - tx, wx_1 = tx, accumulated from zero in (p, q, r) order.  An
 * index out of range gives NaN at that node.  Replaces _trilinear
 * (initflow.py:31-44) of initialize_flow_interpolated (initflow.py:47-120,
 * the restart onto another grid).  One launch on the stream, on device
 * `device`. */
int slc_interp_trilinear(int device, const double* src, int sx, int sy, int sz, const int* ix,
                         const double* tx, int nxq, const int* iy, const double* ty, int nyq,
                         const int* iz, const double* tz, int nzq, double* out, void* stream);

/* Exact bulk-flux constraint: u += U_bulk - ub(u) on the interior, ghosts of
 * u refreshed.  out (4 device doubles): [0] correction, [1] forcing =
 * correction/dt, [2] bulk velocity after the shift, [3] before it.
 * Replaces SLChannelFlow._apply_bulk_forcing (solver.py:694-723). */
int slc_bulk_force(slc_plan* plan, double* u, double* v, double* w, double U_bulk, double dt,
                   double* out, void* stream);

/* slc_bulk_force without the plan's workspace: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc):
 * the call allocates nothing.  Its partial sums (at most 2048 doubles, the
 * most any grid needs) live in a static device array of the library, one per
 * GPU, held until the process ends.  Calls that overlap on one GPU share that
 * array and race: order them on one stream.  Same kernels, same results as
 * slc_bulk_force. */
int slc_bulk_force_noalloc(slc_plan* plan, double* u, double* v, double* w, double U_bulk,
                           double dt, double* out, void* stream);

/* Change the plan's kinematic viscosity nu (= 1/Re_b).  No counterpart in
 * the reference, where nu is fixed when the solver is built (solver.py:
 * self.nu = 1/Re from the config): lets a caller step runs of several Re_b
 * on one plan (a serving loop keeps one solver per grid).  The z-solve
 * factors, which hold nu, are rebuilt at their next use. */
int slc_set_viscosity(slc_plan* plan, double nu);

/* Options.
 * SLC_OPT_GATHER_TILE (no counterpart in the reference): -1 = direct
 * global-load kernels; 0 = fused LDS-tiled step kernels (4x4x16 arrival
 * points per tile, default); 1..3 = older standalone tiled gathers for
 * slc_gather (1: 2x8x32, 2: 2x4x64, 3: 4x4x32).
 * SLC_OPT_FIELD_STORAGE: 64 (default) or 32.  At 32 the gathers of
 * slc_advect, slc_bdf1_predict and slc_bdf2_predict read float copies of the
 * advected fields (made by the library in the same pass that forms U* or in
 * one conversion pass), staged as float in LDS; Lagrange weights,
 * accumulation, feet, the explicit xy-diffusion, z-solves and projection stay
 * FP64.  It is the FP32-storage reading of the reference's "fp32_accum64"
 * interp_dtype (semilag.py:340-356, semilag_triton.py:320-342), whose Triton
 * gather stores and accumulates in fp32.  Selecting 32 allocates the float
 * copies (2 x 4 B per field node) once.  Applies when
 * SLC_OPT_GATHER_TILE >= 0; slc_gather always reads FP64.
 * SLC_OPT_BDF2_FUSED (no counterpart in the reference): 1 (default) = one
 * LDS-tiled pass per component forms both departure feet and both gathers
 * of slc_bdf2_predict; 0 = a far pass, then a near pass reading the near
 * feet and the far values back from device memory.  Both give the same
 * doubles; applies when SLC_OPT_GATHER_TILE >= 0.
 * SLC_OPT_BDF2_XY_RHS: the explicit xy term of slc_bdf2_predict, 0 (default)
 * = 2R^n - R^{n-1} ('extrap'), 1 = R^n ('lagged'); sl.bdf2_xy_rhs,
 * solver.py:262-266, 848-854. */
#define SLC_OPT_GATHER_TILE 1
#define SLC_OPT_FIELD_STORAGE 2
#define SLC_OPT_BDF2_FUSED 3
#define SLC_OPT_BDF2_XY_RHS 4
int slc_set_option(slc_plan* plan, int option, int value);

/* Opt-in launch timing (no counterpart in the reference, whose benchmarks
 * time whole steps with perf_counter + cuda.synchronize,
 * benchmarks/bench_step.py:101-109): while enabled, every labelled launch
 * is bracketed by hipEvents on its stream.  Does not synchronise. */
int slc_profile_enable(slc_plan* plan, int on);

/* Synchronise on the recorded events and write one line per label,
 * "label count total_ms\n", into buf (NUL-terminated); clears the records. */
int slc_profile_report(slc_plan* plan, char* buf, int64_t len);

#ifdef __cplusplus
}
#endif

#endif /* SLC_H */
