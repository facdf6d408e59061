// slc_internal.h -- host-side declarations shared by the library's sources.
#pragma once

#include <hip/hip_runtime.h>

#include "slc_device.h"

namespace slc {

// Poisson coefficient tables (projection.py:47-103, on the fly per mode).
struct PoissonTabs {
    int nx, ny, nz, nky;
    const double* kxsq;   // nx      modified wavenumber squared
    const double* kysq;   // nky
    const double* cl;     // nz      1/(dz_c[k]   dz_f[k])
    const double* cr;     // nz      1/(dz_c[k+1] dz_f[k])
    const double* cs;     // nz      cl + cr
};

// semi-Lagrangian kernels (slc_sl.hip)
// layout 0: ghost-shaped trajectory fields; 1: packed node buffers.  ncl
// non-null: z clamped and the clamped feet counted.
hipError_t run_departure(const Grid& g, int comp, const TrajVel& V, int layout, double dt,
                         double* xd, double* yd, double* zd, unsigned long long* ncl,
                         hipStream_t s);
hipError_t run_gather(const Grid& g, int comp, const double* F, long fsx, long fsy, int fox,
                      int foy, const double* xd, const double* yd, const double* zd, double* out,
                      unsigned long long* ncl, hipStream_t s);
hipError_t run_gather_tiled(const Grid& g, int comp, int shape, const double* F,
                            const double* xd, const double* yd, const double* zd, double* out,
                            unsigned long long* ncl, hipStream_t s);
hipError_t run_locate(const Grid& g, int comp, const double* z, long n, int* exact, int* fast,
                      hipStream_t s);
hipError_t run_stencil(const Grid& g, int comp, const double* xd, const double* yd,
                       const double* zd, int* ix, int* iy, int* k0, unsigned long long* ncl,
                       hipStream_t s);
hipError_t run_advect(const Grid& g, int comp, const TrajVel& V, const double* F, double dt,
                      double* out, unsigned long long* ncl, hipStream_t s);
hipError_t run_bdf1(const Grid& g, int comp, const TrajVel& V, const double* F, double dt,
                    double* out, unsigned long long* ncl, hipStream_t s);
hipError_t run_bdf2(const Grid& g, int comp, int pass, const TrajVel& V, const double* Fn,
                    const double* Fm, double dt, double* out, unsigned long long* ncl,
                    hipStream_t s);

// Per-component arrival tables: x/dx (nx), y/dy (ny), and per arrival z
// (nzi) 2 * its face cell + (z >= the cell's centre split).
struct ArrivalTabs {
    const double* xdx;
    const double* ydy;
    const int* csel;
};

// LDS-tiled step kernels (slc_step.hip); mode: 0 advect, 1 bdf1, 2 far,
// 3 near, 4 gather (feet given), 5 far + near in one pass (no feet arrays)
// Fn32 / Fm32: float copies of F / Fm to gather from (FP32 storage), or null.
// defer (mode 4 only): (tiles of the component) unsigned of device scratch,
// one flag per tile; when given, tiles whose box exceeds the LDS budget are
// gathered by a second pass with a larger budget instead of by direct loads.
hipError_t run_tile_step(const Grid& g, int comp, int mode, const ArrivalTabs& A,
                         const TrajVel& V, const double* F, const double* Fm, const float* Fn32,
                         const float* Fm32, double dt, double* out, double* fx, double* fy,
                         double* fz, unsigned long long* ncl, hipStream_t s,
                         unsigned* defer = nullptr);
// Tiles of the gather at given feet for component comp (the defer flags' count).
long gather_tiles(const Grid& g, int comp);
hipError_t run_arrival_tabs(const Grid& g, int comp, double* xdx, double* ydy, int* cell,
                            hipStream_t s);
hipError_t run_ustar(const Grid& g, double* const o[3], const double* const a[3],
                     const double* const b[3], float* const fa[3], float* const fb[3],
                     hipStream_t s);
hipError_t run_to_f32(const Grid& g, float* const o[3], const double* const a[3],
                      hipStream_t s);

// field operators (slc_ops.hip)
hipError_t run_fill_u64(unsigned long long* p, long n, unsigned long long v, hipStream_t s);
hipError_t run_fill_f64(double* p, long n, double v, hipStream_t s);
hipError_t run_bc(const Grid& g, double* u, double* v, double* w, hipStream_t s);
hipError_t run_bc_p(const Grid& g, double* p, hipStream_t s);
hipError_t run_zfactor(const Grid& g, int comp, double dt, double theta, const double* dz_c,
                       const double* dz_f, double* tab, hipStream_t s);
hipError_t run_zsolve(const Grid& g, int comp, double* f, double dt, double theta,
                      const double* dz_c, const double* dz_f, double* tab, bool refactor,
                      hipStream_t s);
hipError_t run_div_planes(const Grid& g, const double* u, const double* v, const double* w,
                          const double* dz_f, double scale, double* planes, hipStream_t s);
hipError_t run_div_field(const Grid& g, const double* u, const double* v, const double* w,
                         const double* dz_f, double* div, hipStream_t s);
hipError_t run_field_to_planes(const Grid& g, const double* f, double scale, double* planes,
                               hipStream_t s);
hipError_t run_poisson_thomas(const PoissonTabs& t, double* dhat, double* cp, hipStream_t s);
hipError_t run_project_planes(const Grid& g, double* u, double* v, double* w,
                              const double* planes, double inv_n, double dt,
                              const double* dz_c, double* p, hipStream_t s);
hipError_t run_project_p(const Grid& g, double* u, double* v, double* w, const double* p,
                         double dt, const double* dz_c, hipStream_t s);
hipError_t run_bulk(const Grid& g, double* u, double* v, double* w, const double* cv,
                    double total_volume, double U_bulk, double dt, int shift, double* partials,
                    int max_partials, double* scal, hipStream_t s);
hipError_t run_bulk_velocity(const Grid& g, const double* u, const double* cv,
                             double total_volume, double* partials, int max_partials,
                             double* out, hipStream_t s);
hipError_t run_axpby(double* out, const double* a, const double* b, double alpha, double beta,
                     long n, hipStream_t s);
int bulk_partials_needed(const Grid& g);

}  // namespace slc
