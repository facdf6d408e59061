"""Ensemble: many independent channel runs advanced together (include/slc_ens.h).

Each member is a full SL-BDF2 channel with its own grid size, Reynolds number
and time step; one ``step()`` advances every member by one step with one
batched launch per substep.  Members can be added and removed between steps;
a new member takes the BDF1 bootstrap on its first step while the others take
BDF2 steps in the same launches.  A member's trajectory is the one the single
run (``SLChannelFlow.step_sl_bdf2``, solver.py:813-1018) computes from the
same state: the batched kernels run the single-run device code per member.

The reference steps one run per process (solver.py:1280-1500); this class has
no counterpart there.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from .grid import generate_grid
from .plan import field_shapes

OPT_TILE, OPT_GRID, OPT_GRAPH, OPT_BDF2_FUSED, OPT_POOL = 1, 2, 3, 4, 5
N_STATS = 5  # u_bulk, forcing, clamped feet, steps, u_bulk before the shift


def _config_grid(cfg):
    g, dom, flow = cfg["grid"], cfg["domain"], cfg["flow"]
    sl = cfg.get("sl", {})
    bc = cfg.get("boundary_conditions", {})
    if int(sl.get("traj_interp_order", 2)) != 2:
        raise ValueError("ensemble members trace trilinear trajectories (traj_interp_order 2); "
                         "tricubic ones are a single-run option of SLChannelFlow")
    if sl.get("bdf2_pressure", "noninc") != "noninc" or sl.get("bdf2_xy_rhs", "extrap") != "extrap":
        raise ValueError("ensemble members run the production BDF2 options (bdf2_pressure "
                         "'noninc', bdf2_xy_rhs 'extrap'); 'inc' / 'lagged' are single-run "
                         "options of SLChannelFlow")
    return dict(
        nx=int(g["nx"]), ny=int(g["ny"]), nz=int(g["nz"]),
        Lx=float(dom["Lx"]), Ly=float(dom["Ly"]), Lz=float(dom["Lz"]),
        stretching=dom.get("stretching_type", "symmetric"),
        gamma=float(flow["gamma"]), nu=1.0 / float(flow["Re"]),
        U_bulk=float(flow.get("U_bulk", 1.0)),
        order=int(sl.get("interp_order", 4)),
        traj_order=int(sl.get("traj_interp_order", 2)),
        n_iters=int(sl.get("n_traj_iters", 2)),
        top=bc.get("top_wall", {}).get("type", "dirichlet"),
    )


def grid_desc(cfg):
    """slc_grid_desc of a SLChannelFlow-style config dict (plus the z tables it
    points to, which the caller keeps alive for the call)."""
    c = _config_grid(cfg)
    z_f, z_c, _, _ = generate_grid(c["gamma"], c["nz"], c["Lz"], stretching_type=c["stretching"])
    zf = np.ascontiguousarray(z_f.numpy(), dtype=np.float64)
    zc = np.ascontiguousarray(z_c.numpy(), dtype=np.float64)
    tanh_g = torch.tanh(torch.tensor(c["gamma"], dtype=torch.float64)).item()
    P = ctypes.POINTER(ctypes.c_double)
    desc = _lib.GridDesc(c["nx"], c["ny"], c["nz"], c["Lx"], c["Ly"], c["Lz"], c["gamma"], tanh_g,
                         0 if c["stretching"] == "symmetric" else 1, c["order"], c["traj_order"],
                         c["n_iters"], 0 if c["top"] == "dirichlet" else 1, c["nu"],
                         zf.ctypes.data_as(P), zc.ctypes.data_as(P))
    return desc, (zf, zc), c


def member_bytes(cfg, storage=64):
    """Device bytes one member of this config holds (admission by footprint)."""
    lib = _lib.load()
    desc, keep, _ = grid_desc(cfg)
    out = ctypes.c_int64()
    _lib.check(lib.slc_ens_member_bytes(ctypes.byref(desc), int(storage), ctypes.byref(out)),
               "slc_ens_member_bytes")
    del keep
    return out.value


class Ensemble:
    """Batched ensemble of independent channels on one GPU."""

    def __init__(self, order=6, storage=64, device=None, graph=False):
        if not torch.cuda.is_available():
            raise RuntimeError("the HIP layer needs a ROCm GPU (torch.cuda.is_available() is False)")
        self._lib = _lib.load()
        dev = torch.device("cuda") if device is None else torch.device(device)
        if dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self.order = int(order)
        self.storage = int(storage)
        h = ctypes.c_void_p()
        with torch.cuda.device(dev):
            _lib.check(self._lib.slc_ens_create(dev.index, self.order, self.storage,
                                                ctypes.byref(h)), "slc_ens_create")
        self._h = h
        self.members = {}  # id -> dict(nx, ny, nz, shapes, dt, config)
        if graph:
            self.set_option(OPT_GRAPH, 1)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.slc_ens_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    # -- membership ----------------------------------------------------------

    def add(self, config, dt, fields=None):
        """Add a member for a SLChannelFlow-style config dict; fields =
        (u, v, w[, p]) ghost-shaped float64 tensors (copied), or None for
        zero fields.  Returns the member id."""
        desc, keep, c = grid_desc(config)
        if c["order"] != self.order:
            raise ValueError(f"member interp_order {c['order']} != ensemble order {self.order}")
        mid = ctypes.c_int()
        with torch.cuda.device(self.device):
            _lib.check(self._lib.slc_ens_add(self._h, ctypes.byref(desc), float(dt), c["U_bulk"],
                                             ctypes.byref(mid)), "slc_ens_add")
        del keep
        self.members[mid.value] = dict(nx=c["nx"], ny=c["ny"], nz=c["nz"], dt=float(dt),
                                       shapes=field_shapes(c["nx"], c["ny"], c["nz"]),
                                       config=config)
        if fields is not None:
            self.set_fields(mid.value, *fields)
        return mid.value

    def remove(self, mid):
        _lib.check(self._lib.slc_ens_remove(self._h, int(mid)), "slc_ens_remove")
        del self.members[mid]

    def __len__(self):
        return len(self.members)

    def bytes(self):
        out = ctypes.c_int64()
        _lib.check(self._lib.slc_ens_bytes(self._h, ctypes.byref(out)), "slc_ens_bytes")
        return out.value

    def trim_pool(self, keep_bytes=-1):
        """Free released member arenas (oldest first) until the pool holds at
        most keep_bytes (-1: free nothing); returns the bytes it still holds."""
        out = ctypes.c_int64()
        with torch.cuda.device(self.device):
            _lib.check(self._lib.slc_ens_trim_pool(self._h, int(keep_bytes), ctypes.byref(out)),
                       "slc_ens_trim_pool")
        return out.value

    # -- fields ----------------------------------------------------------------

    def _checked(self, mid, name, t):
        shape = self.members[mid]["shapes"][name]
        if not isinstance(t, torch.Tensor) or t.device != self.device or t.dtype != torch.float64:
            raise ValueError(f"{name}: float64 tensor on {self.device} required")
        if tuple(t.shape) != tuple(shape) or not t.is_contiguous():
            raise ValueError(f"{name}: contiguous shape {shape} required, got {tuple(t.shape)}")
        return ctypes.c_void_p(t.data_ptr())

    def set_fields(self, mid, u, v, w, p=None):
        """Copy fields in; the member's next step is a BDF1 bootstrap."""
        args = [self._checked(mid, n, t) for n, t in (("u", u), ("v", v), ("w", w))]
        pp = self._checked(mid, "p", p) if p is not None else ctypes.c_void_p(None)
        _lib.check(self._lib.slc_ens_set_fields(self._h, int(mid), *args, pp, self._stream()),
                   "slc_ens_set_fields")

    def get_fields(self, mid):
        """New tensors holding the member's current u, v, w, p."""
        sh = self.members[mid]["shapes"]
        out = {k: torch.empty(sh[k], dtype=torch.float64, device=self.device) for k in "uvwp"}
        _lib.check(self._lib.slc_ens_get_fields(
            self._h, int(mid), *[ctypes.c_void_p(out[k].data_ptr()) for k in "uvwp"],
            self._stream()), "slc_ens_get_fields")
        return out

    def stats(self, mid):
        """Device tensor: u_bulk, forcing, clamped feet, steps, u_bulk before shift."""
        out = torch.empty(N_STATS, dtype=torch.float64, device=self.device)
        _lib.check(self._lib.slc_ens_stats(self._h, int(mid), ctypes.c_void_p(out.data_ptr()),
                                           self._stream()), "slc_ens_stats")
        return out

    def diagnostics(self, mid, out=None):
        """Device tensor [max|div|, u_tau, CFL, dti, non-finite entries] of the
        member's current fields (no host synchronisation)."""
        if out is None:
            out = torch.empty(5, dtype=torch.float64, device=self.device)
        _lib.check(self._lib.slc_ens_diagnostics(self._h, int(mid), ctypes.c_void_p(out.data_ptr()),
                                                 self._stream()), "slc_ens_diagnostics")
        return out

    # -- stepping ------------------------------------------------------------

    def step(self):
        _lib.check(self._lib.slc_ens_step(self._h, self._stream()), "slc_ens_step")

    def set_option(self, option, value):
        _lib.check(self._lib.slc_ens_set_option(self._h, int(option), int(value)),
                   "slc_ens_set_option")

    def get_option(self, option):
        v = ctypes.c_int()
        _lib.check(self._lib.slc_ens_get_option(self._h, int(option), ctypes.byref(v)),
                   "slc_ens_get_option")
        return v.value

    def autotune(self, reps=3):
        """Time the SL passes for every tile shape and the other substeps for
        every grid mode; keep the fastest.  Returns ({shape: ms}, {grid: ms})."""
        buf = ctypes.create_string_buffer(1024)
        _lib.check(self._lib.slc_ens_autotune(self._h, int(reps), self._stream(), buf, 1024),
                   "slc_ens_autotune")
        shapes, grids = {}, {}
        for line in buf.value.decode().splitlines():
            s, ms = line.split()
            if s.startswith("grid"):
                grids[int(s[4:])] = float(ms)
            else:
                shapes[int(s)] = float(ms)
        return shapes, grids


_INIT_CACHE = {}


def initial_fields(config, seed, urms=0.1, kmin=1.0, kmax=8.0, spectrum="flat", device=None):
    """Seeded solenoidal initial state of one member (SLChannelFlow.initialize:
    Poiseuille + perturbation, boundary-filled and projected), as device
    tensors (u, v, w, p).  One single-run solver per grid is cached."""
    from .solver import SLChannelFlow

    c = _config_grid(config)
    key = (c["nx"], c["ny"], c["nz"], c["Lx"], c["Ly"], c["Lz"], c["gamma"], c["stretching"],
           c["top"], c["order"], str(device))
    flow = _INIT_CACHE.get(key)
    if flow is None:
        flow = SLChannelFlow(config=config, device=device)
        _INIT_CACHE[key] = flow
    flow.U_bulk = c["U_bulk"]
    flow.initialize(seed=seed, urms=urms, kmin=kmin, kmax=kmax, spectrum=spectrum)
    return flow.u.clone(), flow.v.clone(), flow.w.clone(), flow.p.clone()
