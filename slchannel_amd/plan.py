"""HipPlan: the Python face of one slc_plan (libslc_hip.so).

Owns the native plan for one grid and exposes the step's operators on torch
float64 CUDA(HIP) tensors.  Every argument is validated on the host (device,
dtype, contiguity, exact shape) before any kernel is launched, so a kernel
never sees a buffer of the wrong size.  All work goes to torch's current
stream of the plan's device; nothing synchronises the host.
"""

import ctypes

import numpy as np
import torch

from . import _lib

COMPS = "uvw"


def field_shapes(nx, ny, nz):
    """Ghost-shaped staggered array shapes (solver.py conventions)."""
    return {
        "u": (nx + 1, ny + 2, nz + 2),
        "v": (nx + 2, ny + 1, nz + 2),
        "w": (nx + 2, ny + 2, nz + 1),
        "p": (nx + 2, ny + 2, nz + 2),
    }


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(None)


class HipPlan:
    """One native plan: grid tables, Poisson tables, rocFFT plans, work buffers."""

    def __init__(self, nx, ny, nz, Lx, Ly, Lz, z_f, z_c, gamma, stretching_type="symmetric",
                 order=6, traj_order=2, n_traj_iters=2, top_wall_bc_type="dirichlet", nu=0.0,
                 tanh_g=None, device=None):
        if not torch.cuda.is_available():
            raise RuntimeError("the HIP layer needs a ROCm GPU (torch.cuda.is_available() is False)")
        lib = _lib.load()
        dev = torch.device("cuda") if device is None else torch.device(device)
        if dev.type != "cuda":
            raise ValueError(f"HipPlan needs a cuda (HIP) device, got {dev}")
        if dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.Lx, self.Ly, self.Lz = float(Lx), float(Ly), float(Lz)
        self.order = int(order)
        self.nu = float(nu)
        self.shapes = field_shapes(self.nx, self.ny, self.nz)
        zf = np.ascontiguousarray(torch.as_tensor(z_f).detach().cpu().numpy(), dtype=np.float64)
        zc = np.ascontiguousarray(torch.as_tensor(z_c).detach().cpu().numpy(), dtype=np.float64)
        if zf.shape != (self.nz + 1,) or zc.shape != (self.nz + 2,):
            raise ValueError(f"z_f must have nz+1 and z_c nz+2 entries, got {zf.shape} {zc.shape}")
        if tanh_g is None:
            # semilag.py:240-242 evaluates tanh(gamma) as a float64 torch scalar
            tanh_g = torch.tanh(torch.tensor(float(gamma), dtype=torch.float64)).item()
        self.tanh_g = float(tanh_g)
        self._zf, self._zc = zf, zc  # keep alive for the create call
        desc = _lib.GridDesc(
            self.nx, self.ny, self.nz, self.Lx, self.Ly, self.Lz, float(gamma), self.tanh_g,
            0 if stretching_type == "symmetric" else 1, self.order, int(traj_order),
            int(n_traj_iters), 0 if top_wall_bc_type == "dirichlet" else 1, self.nu,
            zf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            zc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
        )
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(lib.slc_plan_create(ctypes.byref(desc), self.device.index or 0,
                                           ctypes.byref(handle)), "slc_plan_create")
        self._h = handle
        self._lib = lib
        eps_z = 1e-12 * self.Lz
        self.z_lo = float(zf[0] + eps_z)
        self.z_hi = float(zf[-1] - eps_z)

    # -- housekeeping -------------------------------------------------------

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.slc_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def workspace_bytes(self):
        out = ctypes.c_int64()
        _lib.check(self._lib.slc_plan_workspace_bytes(self._h, ctypes.byref(out)),
                   "slc_plan_workspace_bytes")
        return out.value

    OPT_GATHER_TILE = 1
    OPT_FIELD_STORAGE = 2
    OPT_BDF2_FUSED = 3
    OPT_BDF2_XY_RHS = 4

    def set_option(self, option, value):
        """Tuning knob (include/slc.h: SLC_OPT_*)."""
        _lib.check(self._lib.slc_set_option(self._h, int(option), int(value)), "slc_set_option")

    def profile(self, on=True):
        """Bracket every labelled launch with hipEvents (see profile_report)."""
        _lib.check(self._lib.slc_profile_enable(self._h, 1 if on else 0), "slc_profile_enable")

    def profile_report(self):
        """{label: (launches, total_ms)} since the last report (synchronises)."""
        buf = ctypes.create_string_buffer(1 << 16)
        _lib.check(self._lib.slc_profile_report(self._h, buf, len(buf)), "slc_profile_report")
        out = {}
        for line in buf.value.decode().splitlines():
            label, n, ms = line.split()
            out[label] = (int(n), float(ms))
        return out

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _chk(self, t, shape, name, dtype=torch.float64):
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{name}: expected a torch.Tensor")
        if t.device != self.device:
            raise ValueError(f"{name}: on {t.device}, plan is on {self.device}")
        if t.dtype != dtype:
            raise TypeError(f"{name}: dtype {t.dtype}, expected {dtype}")
        if not t.is_contiguous():
            raise ValueError(f"{name}: must be contiguous")
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"{name}: shape {tuple(t.shape)}, expected {tuple(shape)}")
        return t

    def _counter(self, nclamp):
        if nclamp is None:
            return ctypes.c_void_p(None)
        self._chk(nclamp, tuple(nclamp.shape), "nclamp", torch.int64)
        if nclamp.numel() < 1:
            raise ValueError("nclamp: needs one element")
        return _ptr(nclamp)

    def _uvw(self, u, v, w, tag=""):
        s = self.shapes
        return (self._chk(u, s["u"], "u" + tag), self._chk(v, s["v"], "v" + tag),
                self._chk(w, s["w"], "w" + tag))

    def n_arrival(self, comp):
        return self.nx * self.ny * (self.nz - 1 if comp == "w" else self.nz)

    def arrival_shape(self, comp):
        return (self.nx, self.ny, self.nz - 1 if comp == "w" else self.nz)

    # -- operators ----------------------------------------------------------

    def apply_bc(self, u=None, v=None, w=None):
        s = self.shapes
        for t, c in ((u, "u"), (v, "v"), (w, "w")):
            if t is not None:
                self._chk(t, s[c], c)
        _lib.check(self._lib.slc_apply_bc(self._h, _ptr(u), _ptr(v), _ptr(w), self._stream()),
                   "slc_apply_bc")

    def node_shape(self, comp):
        """Packed node buffer of a component, (nx, ny, NZ) (semilag.py:363-372)."""
        return (self.nx, self.ny, self.nz + 1 if comp == "w" else self.nz + 2)

    def departure(self, comp, um, vm, wm, dt, xd, yd, zd, nclamp=None, layout=0):
        """Feet of `comp` (include/slc.h slc_departure).  layout 0: ghost-shaped
        trajectory fields; 1: packed node buffers.  With nclamp, z is clamped
        and the clamped feet are counted."""
        if layout == 0:
            self._uvw(um, vm, wm, "_mid")
        elif layout == 1:
            for t, c in ((um, "u"), (vm, "v"), (wm, "w")):
                self._chk(t, self.node_shape(c), f"mbuf[{c}]")
        else:
            raise ValueError(f"layout must be 0 or 1, got {layout}")
        n = self.n_arrival(comp)
        for t, nm in ((xd, "xd"), (yd, "yd"), (zd, "zd")):
            self._chk(t, (n,), nm)
        _lib.check(self._lib.slc_departure(self._h, COMPS.index(comp), _ptr(um), _ptr(vm),
                                           _ptr(wm), int(layout), float(dt), _ptr(xd), _ptr(yd),
                                           _ptr(zd), self._counter(nclamp), self._stream()),
                   "slc_departure")

    def gather(self, comp, field, xd, yd, zd, out, nclamp=None, layout=0):
        if layout == 0:
            self._chk(field, self.shapes[comp], "field")
        else:
            self._chk(field, self.node_shape(comp), "node buffer")
        n = self.n_arrival(comp)
        for t, nm in ((xd, "xd"), (yd, "yd"), (zd, "zd"), (out, "out")):
            self._chk(t, (n,), nm)
        _lib.check(self._lib.slc_gather(self._h, COMPS.index(comp), _ptr(field), int(layout),
                                        _ptr(xd), _ptr(yd), _ptr(zd), _ptr(out),
                                        self._counter(nclamp), self._stream()), "slc_gather")

    def stencil(self, comp, xd, yd, zd, ix, iy, k0, nclamp=None):
        n = self.n_arrival(comp)
        for t, nm in ((xd, "xd"), (yd, "yd"), (zd, "zd")):
            self._chk(t, (n,), nm)
        for t, nm in ((ix, "ix"), (iy, "iy"), (k0, "k0")):
            self._chk(t, (n,), nm, torch.int32)
        _lib.check(self._lib.slc_stencil(self._h, COMPS.index(comp), _ptr(xd), _ptr(yd),
                                         _ptr(zd), _ptr(ix), _ptr(iy), _ptr(k0),
                                         self._counter(nclamp), self._stream()), "slc_stencil")

    def locate_z(self, comp, z, exact, fast):
        """Node-interval index m0 of z (semilag.py:413-420): analytic and fast paths."""
        n = z.numel()
        self._chk(z, (n,), "z")
        self._chk(exact, (n,), "exact", torch.int32)
        self._chk(fast, (n,), "fast", torch.int32)
        _lib.check(self._lib.slc_locate_z(self._h, COMPS.index(comp), _ptr(z), n, _ptr(exact),
                                          _ptr(fast), self._stream()), "slc_locate_z")

    def advect(self, u, v, w, um, vm, wm, dt, ustar, vstar, wstar, nclamp=None):
        self._uvw(u, v, w)
        self._uvw(um, vm, wm, "_mid")
        self._uvw(ustar, vstar, wstar, "star")
        _lib.check(self._lib.slc_advect(self._h, _ptr(u), _ptr(v), _ptr(w), _ptr(um), _ptr(vm),
                                        _ptr(wm), float(dt), _ptr(ustar), _ptr(vstar),
                                        _ptr(wstar), self._counter(nclamp), self._stream()),
                   "slc_advect")

    def bdf1_predict(self, u, v, w, dt, ustar, vstar, wstar, nclamp=None):
        self._uvw(u, v, w)
        self._uvw(ustar, vstar, wstar, "star")
        _lib.check(self._lib.slc_bdf1_predict(self._h, _ptr(u), _ptr(v), _ptr(w), float(dt),
                                              _ptr(ustar), _ptr(vstar), _ptr(wstar),
                                              self._counter(nclamp), self._stream()),
                   "slc_bdf1_predict")

    def bdf2_predict(self, u, v, w, unm1, vnm1, wnm1, dt, ustar, vstar, wstar, nclamp=None):
        self._uvw(u, v, w)
        self._uvw(unm1, vnm1, wnm1, "_nm1")
        self._uvw(ustar, vstar, wstar, "star")
        _lib.check(self._lib.slc_bdf2_predict(self._h, _ptr(u), _ptr(v), _ptr(w), _ptr(unm1),
                                              _ptr(vnm1), _ptr(wnm1), float(dt), _ptr(ustar),
                                              _ptr(vstar), _ptr(wstar), self._counter(nclamp),
                                              self._stream()), "slc_bdf2_predict")

    def zsolve(self, comp, field, dt, theta=1.0):
        self._chk(field, self.shapes[comp], comp)
        _lib.check(self._lib.slc_zsolve(self._h, COMPS.index(comp), _ptr(field), float(dt),
                                        float(theta), self._stream()), "slc_zsolve")

    def divergence(self, u, v, w, out):
        self._uvw(u, v, w)
        self._chk(out, (self.nx, self.ny, self.nz), "div")
        _lib.check(self._lib.slc_divergence(self._h, _ptr(u), _ptr(v), _ptr(w), _ptr(out),
                                            self._stream()), "slc_divergence")

    def poisson(self, div, p, rhs_scale=1.0):
        self._chk(div, (self.nx, self.ny, self.nz), "div")
        self._chk(p, self.shapes["p"], "p")
        _lib.check(self._lib.slc_poisson(self._h, _ptr(div), float(rhs_scale), _ptr(p),
                                         self._stream()), "slc_poisson")

    def project(self, u, v, w, p, dt):
        self._uvw(u, v, w)
        self._chk(p, self.shapes["p"], "p")
        _lib.check(self._lib.slc_project(self._h, _ptr(u), _ptr(v), _ptr(w), _ptr(p), float(dt),
                                         self._stream()), "slc_project")

    def pressure_combine(self, out, a, x, b, y):
        """out = a x + b y over pressure-shaped arrays (include/slc.h)."""
        for t, name in ((out, "out"), (x, "x"), (y, "y")):
            self._chk(t, self.shapes["p"], name)
        _lib.check(self._lib.slc_pressure_combine(self._h, _ptr(out), float(a), _ptr(x), float(b),
                                                  _ptr(y), self._stream()),
                   "slc_pressure_combine")

    def project_step(self, u, v, w, rhs_scale, dt_proj, p=None):
        self._uvw(u, v, w)
        if p is not None:
            self._chk(p, self.shapes["p"], "p")
        _lib.check(self._lib.slc_project_step(self._h, _ptr(u), _ptr(v), _ptr(w),
                                              float(rhs_scale), float(dt_proj), _ptr(p),
                                              self._stream()), "slc_project_step")

    def bulk_velocity(self, u, out):
        self._chk(u, self.shapes["u"], "u")
        self._chk(out, (4,), "out")
        _lib.check(self._lib.slc_bulk_velocity(self._h, _ptr(u), _ptr(out), self._stream()),
                   "slc_bulk_velocity")

    def imex_rhs(self, u, v, w, ru, rv, rw):
        """compute_momentum_rhs_fused_imex into ru, rv, rw (shapes of u, v, w)."""
        self._uvw(u, v, w)
        self._uvw(ru, rv, rw, "_rhs")
        _lib.check(self._lib.slc_imex_rhs(self._h, _ptr(u), _ptr(v), _ptr(w), _ptr(ru), _ptr(rv),
                                          _ptr(rw), self._stream()), "slc_imex_rhs")

    def ab2_update(self, u, v, w, r, r_prev, dt):
        """u, v, w += dt (1.5 r - 0.5 r_prev); += dt r when r_prev is None."""
        self._uvw(u, v, w)
        self._uvw(*r, "_rhs")
        if r_prev is not None:
            self._uvw(*r_prev, "_prev")
        rp = [_ptr(t) for t in r_prev] if r_prev is not None else [_ptr(None)] * 3
        _lib.check(self._lib.slc_ab2_update(self._h, _ptr(u), _ptr(v), _ptr(w),
                                            *[_ptr(t) for t in r], *rp, float(dt),
                                            self._stream()), "slc_ab2_update")

    def set_viscosity(self, nu):
        """slc_set_viscosity: the plan's nu (z-solve factors rebuilt at next use)."""
        _lib.check(self._lib.slc_set_viscosity(self._h, float(nu)), "slc_set_viscosity")
        self.nu = float(nu)

    def init_solenoidal(self, W, L, ab, zt, phase, urms, work, u, v, w):
        """Seeded solenoidal interiors of u, v, w (slc_init_solenoidal): ab, zt,
        phase and work are float64 device tensors laid out as include/slc.h
        describes; ghosts are not written."""
        self._uvw(u, v, w)
        for t, name in ((ab, "ab"), (zt, "zt"), (phase, "phase"), (work, "work")):
            if not (t.is_contiguous() and t.dtype == torch.float64 and t.device == self.device):
                raise ValueError(f"init_solenoidal: {name} must be a contiguous float64 tensor "
                                 f"on {self.device}")
        _lib.check(self._lib.slc_init_solenoidal(
            self._h, int(W), int(L), _ptr(ab), _ptr(zt), _ptr(phase), float(urms), _ptr(work),
            int(work.numel() * 8), _ptr(u), _ptr(v), _ptr(w), self._stream()),
            "slc_init_solenoidal")

    def init_solenoidal_work(self, W):
        """Scratch bytes slc_init_solenoidal needs for W x W bins."""
        return int(self._lib.slc_init_solenoidal_work(self.nx, self.ny, self.nz, int(W)))

    def diagnostics(self, u, v, w, p, dt, out):
        """out (5 device doubles): max|div|, u_tau, CFL = dti*dt, dti, number
        of non-finite entries of u, v, w, p (p may be None)."""
        self._uvw(u, v, w)
        if p is not None:
            self._chk(p, self.shapes["p"], "p")
        self._chk(out, (5,), "out")
        _lib.check(self._lib.slc_diagnostics(self._h, _ptr(u), _ptr(v), _ptr(w),
                                             _ptr(p), float(dt),
                                             _ptr(out), self._stream()), "slc_diagnostics")

    def bulk_force(self, u, v, w, U_bulk, dt, out):
        self._uvw(u, v, w)
        self._chk(out, (4,), "out")
        _lib.check(self._lib.slc_bulk_force(self._h, _ptr(u), _ptr(v), _ptr(w), float(U_bulk),
                                            float(dt), _ptr(out), self._stream()),
                   "slc_bulk_force")
