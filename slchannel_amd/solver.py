"""SLChannelFlow on MI355X: the SL-BDF2 step and the projection around it.

Keeps the reference's step API -- ``step_sl_bdf2(dt) -> (u_bulk, forcing)``
(slchannel/solver.py:813-1018) with the BDF1 bootstrap on the first step and
on any dt change (solver.py:729-811, 834-844), ``apply_bc_uvw`` and the exact
bulk-flux constraint (solver.py:694-723) -- and the YAML keys of the hot path
(grid, domain, flow, boundary_conditions, time.dt, sl.*; docs/CONFIG.md).
Restarts read the reference's ``fields.npz`` layout
(``initialization.field_file``, solver.py:300-347; slchannel_amd.checkpoint)
and re-bootstrap with one BDF1 step; statistics are
slchannel_amd.TurbulenceStats.  The run loop and plotting of the reference
are not part of this layer.

Per step the device does: ghost fill; the fused two-foot BDF2 predictor (far
pass + near pass per component); ghost fill; three Thomas z-solves; ghost fill;
divergence -> rocFFT -> per-mode Thomas -> inverse rocFFT -> gradient
projection -> ghost fill; the deterministic bulk shift.  No host
synchronisation: u_bulk and forcing come back as 0-D device tensors, as in the
reference.
"""

import numpy as np
import torch
import yaml

from . import env
from .checkpoint import initialize_flow_from_file, interpolated_fields, save_flow_fields
from .grid import generate_grid
from .initflow import solenoidal_fields
from .plan import HipPlan, field_shapes
from .semilag import SLAdvector


def _load_config(config_file=None, config=None):
    if config is not None:
        return config
    with open(config_file) as f:
        return yaml.safe_load(f)


class SLChannelFlow:
    """Channel-flow SL-BDF2 stepper on one HIP device.

    ``config_file`` (YAML path) or ``config`` (dict) uses the reference's keys;
    fields start as None until :meth:`set_fields` or :meth:`initialize`.
    """

    def __init__(self, config_file="config.yaml", config=None, device=None, z_tables=None):
        cfg = _load_config(config_file, config)
        self.nx = int(cfg["grid"]["nx"])
        self.ny = int(cfg["grid"]["ny"])
        self.nz = int(cfg["grid"]["nz"])
        dom = cfg["domain"]
        self.Lx, self.Ly, self.Lz = float(dom["Lx"]), float(dom["Ly"]), float(dom["Lz"])
        self.stretching_type = dom.get("stretching_type", "symmetric")
        if self.stretching_type not in ("symmetric", "bottom"):
            raise ValueError(f"Invalid stretching type: {self.stretching_type}. "
                             "slChannel supports 'symmetric' or 'bottom'")
        flow = cfg["flow"]
        self.nu = 1.0 / float(flow["Re"])
        self.U_bulk = float(flow.get("U_bulk", 1.0))
        self.gamma = float(flow["gamma"])
        bc = cfg.get("boundary_conditions", {})
        self.top_wall_bc_type = bc.get("top_wall", {}).get("type", "dirichlet")
        if self.top_wall_bc_type not in ("dirichlet", "neumann"):
            raise ValueError(f"Invalid top wall BC type: {self.top_wall_bc_type}")
        self.dt = cfg.get("time", {}).get("dt", None)
        sl = cfg.get("sl", {})
        self.sl_order = int(sl.get("interp_order", 4))
        self.sl_traj_order = int(sl.get("traj_interp_order", 2))
        self.sl_traj_iters = int(sl.get("n_traj_iters", 2))
        self.sl_interp_dtype = sl.get("interp_dtype", "fp64")
        # BDF2 options (solver.py:256-281): 'noninc' p^{n+1} = phi, or 'inc'
        # on p_ext = 2 p^n - p^{n-1}; xy RHS 2R^n - R^{n-1} ('extrap') or R^n
        self.sl_bdf2_pressure = sl.get("bdf2_pressure", "noninc")
        self.sl_bdf2_xy_rhs = sl.get("bdf2_xy_rhs", "extrap")
        if self.sl_bdf2_pressure not in ("noninc", "inc"):
            raise ValueError(
                f"sl.bdf2_pressure must be 'noninc' or 'inc', got {self.sl_bdf2_pressure}")
        if self.sl_bdf2_xy_rhs not in ("extrap", "lagged"):
            raise ValueError(
                f"sl.bdf2_xy_rhs must be 'extrap' or 'lagged', got {self.sl_bdf2_xy_rhs}")

        if device is None:
            device = torch.device("cuda", torch.cuda.current_device()) \
                if torch.cuda.is_available() else None
        if device is None or torch.device(device).type != "cuda":
            raise RuntimeError("SLChannelFlow (HIP layer) needs a ROCm GPU device")
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        if not env.USE_HIP:
            raise RuntimeError("SLCHANNEL_HIP=0: the HIP layer is disabled")

        if z_tables is None:
            self.z_f, self.z_c, self.dz_f, self.dz_c = generate_grid(
                self.gamma, self.nz, self.Lz, stretching_type=self.stretching_type)
        else:
            # caller-supplied tables (e.g. a grid read from a checkpoint)
            self.z_f = torch.as_tensor(z_tables[0], dtype=torch.float64).cpu()
            self.z_c = torch.as_tensor(z_tables[1], dtype=torch.float64).cpu()
            self.dz_f = self.z_f[1:] - self.z_f[:-1]
            self.dz_c = self.z_c[1:] - self.z_c[:-1]
        self.dx = self.Lx / self.nx
        self.dy = self.Ly / self.ny
        self.total_volume = self.Lx * self.Ly * self.Lz

        self.plan = HipPlan(self.nx, self.ny, self.nz, self.Lx, self.Ly, self.Lz, self.z_f,
                            self.z_c, self.gamma, stretching_type=self.stretching_type,
                            order=self.sl_order, traj_order=self.sl_traj_order,
                            n_traj_iters=self.sl_traj_iters,
                            top_wall_bc_type=self.top_wall_bc_type, nu=self.nu,
                            device=self.device)
        if self.sl_bdf2_xy_rhs == "lagged":
            self.plan.set_option(HipPlan.OPT_BDF2_XY_RHS, 1)
        self.sl = SLAdvector(self.nx, self.ny, self.nz, self.dx, self.dy, self.Lx, self.Ly,
                             self.Lz, self.z_f, self.z_c, self.gamma,
                             stretching_type=self.stretching_type, order=self.sl_order,
                             traj_order=self.sl_traj_order, n_traj_iters=self.sl_traj_iters,
                             top_wall_bc_type=self.top_wall_bc_type,
                             interp_dtype=self.sl_interp_dtype, plan=self.plan)
        self.shapes = field_shapes(self.nx, self.ny, self.nz)
        self.u = self.v = self.w = self.p = None
        self.u_nm1 = self.v_nm1 = self.w_nm1 = None
        self._star = None
        self._dt_prev_step = None
        self._imex_cur = self._imex_prev = None  # IMEX rhs buffers (step_imex)
        # 'inc' pressure: Poisson output phi, full levels p^n / p^{n-1}, p_ext
        self._phi = self._P_curr = self._P_prev = self._p_ext = None
        self.forcing = 0.0
        self.current_step = 0
        self.time = 0.0
        self.initial_step = 0
        self.initial_time = 0.0
        self._nclamp = torch.zeros(1, dtype=torch.int64, device=self.device)
        self._bulk = torch.zeros(4, dtype=torch.float64, device=self.device)

        init = cfg.get("initialization") or {}
        field_file = init.get("field_file")
        if init.get("type", "parabolic") == "interpolate":
            if field_file is None:
                raise ValueError(
                    "initialization.type 'interpolate' requires initialization.field_file")
            self.initialize_interpolated(field_file, source_half=init.get("source_half", "lower"))
        elif field_file is not None:
            self.restart_from(field_file, reset_time=bool(init.get("reset_time", False)))

    # ------------------------------------------------------------------
    # fields
    # ------------------------------------------------------------------

    def _dev(self, a, shape):
        t = torch.as_tensor(a, dtype=torch.float64)
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"field shape {tuple(t.shape)}, expected {tuple(shape)}")
        return t.to(self.device).contiguous().clone()

    def set_fields(self, u, v, w, p=None):
        """Adopt ghost-shaped fields (copied to the device).  Resets histories,
        so the next step is a BDF1 bootstrap (as a restart is)."""
        s = self.shapes
        self.u, self.v, self.w = self._dev(u, s["u"]), self._dev(v, s["v"]), self._dev(w, s["w"])
        self.p = self._dev(p, s["p"]) if p is not None else torch.zeros(
            s["p"], dtype=torch.float64, device=self.device)
        self.u_nm1 = self.v_nm1 = self.w_nm1 = None
        self._dt_prev_step = None
        self._imex_prev = None

    def restart_from(self, field_file, reset_time=False):
        """Restart from a reference-layout checkpoint (solver.py:338-351):
        fields, step and time (unless reset_time) and the forcing value; the
        next step is a BDF1 bootstrap."""
        u, v, w, p, step, t = initialize_flow_from_file(field_file, device=self.device,
                                                        reset_time=reset_time)
        self.set_fields(u, v, w, p)
        self.initial_step = self.current_step = step
        self.initial_time = self.time = t
        try:
            self.forcing = float(np.load(field_file)["forcing"])
        except (KeyError, ValueError):
            self.forcing = 0.0

    def save_checkpoint(self, results_folder, filename="fields.npz", u_tau=0.0):
        """Write the current state in the reference's fields.npz layout
        (utils.py:183-215)."""
        save_flow_fields(self.u, self.v, self.w, self.p, self.z_c, self.z_f, self.Lx, self.Ly,
                         self.current_step, self.time, u_tau, self.forcing, results_folder,
                         filename)

    def initialize(self, seed, urms=0.1, kmin=1.0, kmax=8.0, spectrum="flat"):
        """Seeded solenoidal Poiseuille + perturbation, boundary-filled and
        projected with dt = 1 (solver.py:599-618)."""
        u, v, w = solenoidal_fields(self.nx, self.ny, self.nz, self.Lx, self.Ly, self.Lz,
                                    self.z_f.numpy(), self.z_c.numpy(), seed, urms=urms,
                                    kmin=kmin, kmax=kmax, spectrum=spectrum, U_bulk=self.U_bulk,
                                    plan=self.plan)
        self.set_fields(u, v, w)
        self.apply_bc_uvw()
        self.project_initial()

    def initialize_interpolated(self, field_file, source_half="lower"):
        """Restart onto this grid from a stored field of another grid or box
        (solver.py:351-401, initflow.py:47-120): trilinear interpolation of
        each component at its staggered nodes on the device, boundary fill, u
        rescaled to U_bulk (when the interpolated bulk velocity is not ~0),
        then the initial projection (solver.py:599-618).  Step and time start
        at zero; the next step is a BDF1 bootstrap."""
        u, v, w, p = interpolated_fields(field_file, self.nx, self.ny, self.nz, self.Lx, self.Ly,
                                         self.Lz, self.z_c, self.z_f, self.device,
                                         source_half=source_half)
        self.set_fields(u, v, w, p)
        self.apply_bc_uvw()
        self.plan.bulk_velocity(self.u, self._bulk)
        ub = float(self._bulk[3].item())
        if abs(ub) > 1e-9:
            self.u *= self.U_bulk / ub
        self.project_initial()
        self.forcing = 0.0
        self.initial_step = self.current_step = 0
        self.initial_time = self.time = 0.0

    def project_initial(self):
        """solver.py:599-618: one projection with dt = 1 and a ghost fill."""
        self.plan.project_step(self.u, self.v, self.w, 1.0, 1.0, self.p)

    def set_reynolds(self, Re):
        """Change the bulk Reynolds number of this solver (nu = 1/Re; the
        plan's z-solve factors are rebuilt at their next use).  No
        counterpart in the reference, which fixes nu when the solver is built;
        with set_fields it lets one solver step runs of several Re_b."""
        self.nu = 1.0 / float(Re)
        self.plan.set_viscosity(self.nu)

    def apply_bc_uvw(self):
        """apply_bc_all (solver.py:70-107) on the current fields."""
        self.plan.apply_bc(self.u, self.v, self.w)

    def _ensure_star(self):
        if self._star is None:
            s = self.shapes
            self._star = [torch.zeros(s[c], dtype=torch.float64, device=self.device)
                          for c in "uvw"]

    # ------------------------------------------------------------------
    # step
    # ------------------------------------------------------------------

    def _apply_bulk_forcing(self, dt):
        """solver.py:694-723: exact uniform shift of u onto U_bulk; forcing is
        the correction / dt.  Returns 0-D device tensors."""
        self.plan.bulk_force(self.u, self.v, self.w, self.U_bulk, dt, self._bulk)
        self.forcing = self._bulk[1].clone()
        return self._bulk[2].clone(), self.forcing

    def _implicit_and_project(self, dt_z, rhs_scale, dt_proj, theta=1.0, sl=True):
        """(Ghosts,) z-solves, ghosts, projection; returns the buffer holding
        phi: self.p, or for SL steps with the 'inc' option a separate phi
        buffer (self.p then ends as the full pressure p^{n+1})."""
        plan = self.plan
        # the z-solves read ghost values only for the explicit part of a
        # theta < 1 right-hand side; at theta = 1 the fill the reference makes
        # here (solver.py:766, 949) is overwritten by the one after the solves
        # before anything reads it
        if theta != 1.0:
            plan.apply_bc(self.u, self.v, self.w)
        plan.zsolve("u", self.u, dt_z, theta)
        plan.zsolve("v", self.v, dt_z, theta)
        plan.zsolve("w", self.w, dt_z, theta)
        plan.apply_bc(self.u, self.v, self.w)
        if sl and self.sl_bdf2_pressure == "inc":
            if self._phi is None:
                self._phi = torch.zeros_like(self.p)
            phi = self._phi
        else:
            if self.p is self._P_curr or self.p is self._P_prev:
                self.p = torch.empty_like(self.p)  # keep the 'inc' history intact
            phi = self.p
        plan.project_step(self.u, self.v, self.w, rhs_scale, dt_proj, phi)
        return phi

    def _step_sl_bdf1_bootstrap(self, dt):
        """solver.py:729-811: U* = V^n, one foot at dt, theta = 1 at dt."""
        self._ensure_star()
        if self.u_nm1 is None:
            self.u_nm1, self.v_nm1, self.w_nm1 = self.u.clone(), self.v.clone(), self.w.clone()
        else:
            self.u_nm1.copy_(self.u)
            self.v_nm1.copy_(self.v)
            self.w_nm1.copy_(self.w)
        self._nclamp.zero_()
        us, vs, ws = self._star
        self.plan.bdf1_predict(self.u, self.v, self.w, dt, us, vs, ws, self._nclamp)
        # adopt the predictor buffers; the old field buffers become scratch
        self._star = [self.u, self.v, self.w]
        self.u, self.v, self.w = us, vs, ws
        phi = self._implicit_and_project(dt, dt, dt)
        if self.sl_bdf2_pressure == "inc":
            # solver.py:802-808: p^n = phi, no p^{n-1} yet
            if self._P_curr is None:
                self._P_curr = phi.clone()
            else:
                self._P_curr.copy_(phi)
            self._P_prev = None
            self.p = self._P_curr
        self.sl.n_clamped_last = self._nclamp[0].clone()
        self._dt_prev_step = dt
        return self._apply_bulk_forcing(dt)

    def step_sl_bdf2(self, dt):
        """solver.py:813-1018 (noninc pressure, extrapolated xy RHS)."""
        dt = float(dt)
        self.apply_bc_uvw()
        if (self.u_nm1 is None or self._dt_prev_step is None
                or abs(dt - self._dt_prev_step) > 1e-12 * dt):
            if self._dt_prev_step is not None and self.u_nm1 is not None:
                print(f"[bdf2] dt changed ({self._dt_prev_step:.6g} -> {dt:.6g}): "
                      "BDF1 re-bootstrap", flush=True)
            out = self._step_sl_bdf1_bootstrap(dt)
        else:
            self._ensure_star()
            self._nclamp.zero_()
            us, vs, ws = self._star
            self.plan.bdf2_predict(self.u, self.v, self.w, self.u_nm1, self.v_nm1, self.w_nm1,
                                   dt, us, vs, ws, self._nclamp)
            dt_eff = (2.0 / 3.0) * dt
            p_ext = self._pressure_extrapolation()
            if p_ext is not None:
                # solver.py:918-941: predictor += dt_eff * (-grad p_ext) before
                # the implicit solve; u -= dt_eff grad p_ext is the same doubles
                self.plan.project(us, vs, ws, p_ext, dt_eff)
            # rotate: V^{n-1} <- V^n, V <- predictor, scratch <- old V^{n-1}
            self._star = [self.u_nm1, self.v_nm1, self.w_nm1]
            self.u_nm1, self.v_nm1, self.w_nm1 = self.u, self.v, self.w
            self.u, self.v, self.w = us, vs, ws
            phi = self._implicit_and_project(dt_eff, dt_eff, dt_eff)
            if self.sl_bdf2_pressure == "inc":
                self._pressure_update(p_ext, phi)
            self.sl.n_clamped_last = self._nclamp[0].clone()
            self._dt_prev_step = dt
            out = self._apply_bulk_forcing(dt)
        self.current_step += 1
        self.time += dt
        return out

    def _pressure_extrapolation(self):
        """solver.py:906-917 ('inc'): p_ext = 2 p^n - p^{n-1}, or p^n when
        there is no p^{n-1} (first BDF2 step after a bootstrap)."""
        if self.sl_bdf2_pressure != "inc" or self._P_curr is None:
            return None
        if self._P_prev is None:
            return self._P_curr
        if self._p_ext is None:
            self._p_ext = torch.empty_like(self._P_curr)
        self.plan.pressure_combine(self._p_ext, 2.0, self._P_curr, -1.0, self._P_prev)
        return self._p_ext

    def _pressure_update(self, p_ext, phi):
        """solver.py:999-1012 ('inc'): p^{n+1} = p_ext + phi, p^{n-1} <- p^n."""
        if self._P_curr is None:
            self._P_curr = phi.clone()
        else:
            # the old p^{n-1} buffer is free once p_ext is formed
            nxt = self._P_prev if self._P_prev is not None else torch.empty_like(phi)
            self.plan.pressure_combine(nxt, 1.0, p_ext, 1.0, phi)
            self._P_prev, self._P_curr = self._P_curr, nxt
        self.p = self._P_curr

    # ------------------------------------------------------------------
    # Eulerian IMEX reference scheme (solver.py:1024-1127)
    # ------------------------------------------------------------------

    def compute_momentum_rhs_explicit_imex(self):
        """solver.py:1024-1053: (rhs_u, rhs_v, rhs_w) = diffusion_xy - advection,
        new device tensors."""
        s = self.shapes
        r = [torch.empty(s[c], dtype=torch.float64, device=self.device) for c in "uvw"]
        self.plan.imex_rhs(self.u, self.v, self.w, *r)
        return tuple(r)

    def step_imex(self, dt):
        """solver.py:1055-1127: AB2 explicit advection + xy-diffusion (forward
        Euler while there is no previous rhs), Crank-Nicolson z-diffusion,
        projection with dt, bulk shift.  Returns (u_bulk, forcing) as 0-D
        device tensors."""
        dt = float(dt)
        self.apply_bc_uvw()
        s = self.shapes
        if self._imex_cur is None:
            self._imex_cur = [torch.empty(s[c], dtype=torch.float64, device=self.device)
                              for c in "uvw"]
        self.plan.imex_rhs(self.u, self.v, self.w, *self._imex_cur)
        prev = self._imex_prev
        self.plan.ab2_update(self.u, self.v, self.w, self._imex_cur, prev, dt)
        # the current rhs becomes the previous one (buffers swapped, no copy)
        self._imex_prev, self._imex_cur = self._imex_cur, prev
        self._implicit_and_project(dt, dt, dt, theta=0.5, sl=False)
        out = self._apply_bulk_forcing(dt)
        self.current_step += 1
        self.time += dt
        return out

    # ------------------------------------------------------------------
    # diagnostics (device-side, no host sync unless the caller reads them)
    # ------------------------------------------------------------------

    def divergence(self):
        out = torch.empty(self.nx, self.ny, self.nz, dtype=torch.float64, device=self.device)
        self.plan.divergence(self.u, self.v, self.w, out)
        return out

    def diagnostics(self, dt=None, out=None):
        """Run-loop diagnostics on the device (solver.py:1355-1392, 1436-1479):
        a (5,) float64 device tensor [max|div|, u_tau, CFL, dti, non-finite
        entries of u, v, w, p].  No host synchronisation; ``out`` may be a row
        of a caller-held time-series tensor."""
        if out is None:
            out = torch.empty(5, dtype=torch.float64, device=self.device)
        dt = self.dt if dt is None else dt
        self.plan.diagnostics(self.u, self.v, self.w, self.p, float(dt or 0.0), out)
        return out

    def bulk_velocity(self):
        out = torch.zeros(4, dtype=torch.float64, device=self.device)
        self.plan.bulk_velocity(self.u, out)
        return out[3]
