"""Semi-Lagrangian advection on MI355X: drop-in SLAdvector + HipSL launch wrapper.

``SLAdvector`` keeps the reference's constructor and ``advect`` contract
(slchannel/semilag.py:130-671): same arguments and validation errors, same
ghost-shaped persistent outputs (interiors written, ghosts left alone), the
same ``n_clamped_last`` diagnostic as a 0-D int64 device tensor.  The work
is done by libslc_hip.so: departure feet by the iterated midpoint rule with
trilinear trajectory sampling and the analytic tanh inverse, then the
order-4/6 tensor-product Lagrange gather, all in float64.

``HipSL`` mirrors the reference's plugin object ``TritonSL``
(slchannel/semilag_triton.py:347-455): ``departure(comp, dt) -> n`` and
``gather(comp, buf, n) -> Tensor``, reading the trajectory velocity set by the
last ``advect`` (or ``set_mid``) call.  It is what a maintainer binds in place
of ``TritonSL``.

There is no CPU or eager fallback: constructing either class without a ROCm
device or without the built library raises.
"""

import torch

from . import env
from .plan import HipPlan


class HipSL:
    """Launch wrapper bound to one advector (TritonSL's method shape,
    slchannel/semilag_triton.py:347-455).

    ``adv`` is either this package's ``SLAdvector`` (its plan is reused) or
    the reference's ``slchannel.semilag.SLAdvector`` (a plan is built from its
    grid attributes, as ``TritonSL.__init__`` builds its tables).  The call
    protocol is the reference's (semilag.py:617-666): ``nclamp.zero_()``, then
    per component ``n = departure(comp, dt)`` -- feet of the trajectory
    velocity, z clamped, clamped feet counted into ``nclamp`` -- and
    ``gather(comp, buf, n)`` on a node buffer or a ghost-shaped field, whose
    returned tensor is reused by the next gather of the component.

    The trajectory velocity ``departure`` samples is, in order: the
    ghost-shaped triple of the last ``set_mid`` call, else the advector's
    node buffers ``adv.mbuf`` (semilag.py:604-606, filled by the reference's
    ``advect`` before it calls ``departure``).  Everything is float64.
    """

    def __init__(self, adv):
        self.adv = adv
        plan = getattr(adv, "plan", None)
        if plan is None:
            plan = HipPlan(adv.nx, adv.ny, adv.nz, adv.Lx, adv.Ly, adv.Lz, adv.z_f, adv.z_c,
                           adv.gamma, stretching_type=adv.stretching_type, order=adv.order,
                           traj_order=getattr(adv, "traj_order", 2),
                           n_traj_iters=adv.n_traj_iters,
                           top_wall_bc_type=getattr(adv, "top_wall_bc_type", "dirichlet"),
                           tanh_g=float(adv._tanh_g), device=adv.device)
        self.plan = plan
        dev = plan.device
        n_max = plan.nx * plan.ny * plan.nz
        self.xd = torch.empty(n_max, dtype=torch.float64, device=dev)
        self.yd = torch.empty(n_max, dtype=torch.float64, device=dev)
        self.zd = torch.empty(n_max, dtype=torch.float64, device=dev)
        self.nclamp = torch.zeros(1, dtype=torch.int64, device=dev)
        self.out = {c: torch.empty(plan.arrival_shape(c), dtype=torch.float64, device=dev)
                    for c in "uvw"}
        self._mid = None

    def set_mid(self, u_mid, v_mid, w_mid):
        """Ghost-shaped, BC-consistent trajectory velocity for departure();
        ``set_mid(None, None, None)`` returns to the advector's node buffers."""
        self._mid = None if u_mid is None else (u_mid, v_mid, w_mid)

    def _trajectory(self):
        if self._mid is not None:
            return self._mid, 0
        mbuf = getattr(self.adv, "mbuf", None)
        if mbuf is None:
            raise RuntimeError("HipSL.departure: no trajectory velocity (call set_mid, or bind "
                               "an advector with node buffers mbuf)")
        return (mbuf["u"], mbuf["v"], mbuf["w"]), 1

    def departure(self, comp, dt):
        """Feet of component `comp` for depth dt into self.xd/yd/zd, z clamped
        into [z_lo, z_hi]; the clamped feet are added to self.nclamp.  Returns
        the number of arrival points n."""
        plan = self.plan
        n = plan.n_arrival(comp)
        (um, vm, wm), layout = self._trajectory()
        plan.departure(comp, um, vm, wm, float(dt), self.xd[:n], self.yd[:n], self.zd[:n],
                       nclamp=self.nclamp, layout=layout)
        return n

    def gather(self, comp, buf, n):
        """Interpolate `buf` (ghost-shaped field or (nx, ny, NZ) node buffer) at
        the last feet (clamps were counted by departure()).  Returns the reused
        arrival-shaped output buffer of `comp`."""
        plan = self.plan
        if n != plan.n_arrival(comp):
            raise ValueError(f"HipSL.gather: n={n}, component {comp} has "
                             f"{plan.n_arrival(comp)} arrival points")
        layout = 0 if tuple(buf.shape) == plan.shapes[comp] else 1
        out = self.out[comp].reshape(-1)
        plan.gather(comp, buf, self.xd[:n], self.yd[:n], self.zd[:n], out, None, layout=layout)
        return self.out[comp]


class SLAdvector:
    """Semi-Lagrangian advection of the three staggered velocity components
    (constructor and advect() as slchannel/semilag.py:133-202, 584-602)."""

    def __init__(self, nx, ny, nz, dx, dy, Lx, Ly, Lz, z_f, z_c, gamma,
                 stretching_type="symmetric", order=4, traj_order=2, n_traj_iters=2,
                 top_wall_bc_type="dirichlet", interp_dtype="fp64", device=None, plan=None):
        if order not in (4, 6):
            raise ValueError(f"sl.interp_order must be 4 (tricubic) or 6 (triquintic), got {order}")
        if traj_order not in (2, 4):
            raise ValueError(
                f"sl.traj_interp_order must be 2 (trilinear) or 4 (tricubic), got {traj_order}")
        if stretching_type not in ("symmetric", "bottom"):
            raise ValueError(
                f"SLAdvector supports 'symmetric'/'bottom' stretching, got '{stretching_type}'")
        if interp_dtype not in ("fp64", "fp32_accum64"):
            raise ValueError(
                f"sl.interp_dtype must be 'fp64' or 'fp32_accum64', got {interp_dtype}")
        if nz + 1 < order:
            raise ValueError(f"nz={nz} too small for interpolation order {order}")
        if not env.USE_HIP:
            raise RuntimeError("SLCHANNEL_HIP=0: the HIP semi-Lagrangian layer is disabled")
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dx, self.dy = dx, dy
        self.Lx, self.Ly, self.Lz = Lx, Ly, Lz
        self.gamma = float(gamma)
        self.stretching_type = stretching_type
        self.order = order
        self.traj_order = traj_order
        self.n_traj_iters = n_traj_iters
        self.top_wall_bc_type = top_wall_bc_type
        self.interp_dtype = interp_dtype
        if plan is None:
            plan = HipPlan(nx, ny, nz, Lx, Ly, Lz, z_f, z_c, gamma,
                           stretching_type=stretching_type, order=order, traj_order=traj_order,
                           n_traj_iters=n_traj_iters, top_wall_bc_type=top_wall_bc_type,
                           device=device)
        self.plan = plan
        # fp32_accum64: gathered fields stored as float, weights/accumulation
        # and everything else FP64 (the reference runs the whole interpolation
        # in fp32 there, semilag.py:249-256)
        plan.set_option(HipPlan.OPT_FIELD_STORAGE, 32 if interp_dtype == "fp32_accum64" else 64)
        self.device = plan.device
        self.z_lo, self.z_hi = plan.z_lo, plan.z_hi
        self._tanh_g = plan.tanh_g
        dev = self.device
        self.ustar = torch.zeros(nx + 1, ny + 2, nz + 2, dtype=torch.float64, device=dev)
        self.vstar = torch.zeros(nx + 2, ny + 1, nz + 2, dtype=torch.float64, device=dev)
        self.wstar = torch.zeros(nx + 2, ny + 2, nz + 1, dtype=torch.float64, device=dev)
        self._hip = HipSL(self)
        self.n_clamped_last = torch.zeros((), dtype=torch.int64, device=dev)

    def departure_coords(self, comp, dt_t, u_mid, v_mid, w_mid):
        """Feet (x, y unwrapped; z not clamped) of `comp`, each (nx, ny, nzi)."""
        plan = self.plan
        n = plan.n_arrival(comp)
        hip = self._hip
        plan.departure(comp, u_mid, v_mid, w_mid, float(dt_t), hip.xd[:n], hip.yd[:n],
                       hip.zd[:n])
        shape = plan.arrival_shape(comp)
        return tuple(t[:n].clone().reshape(shape) for t in (hip.xd, hip.yd, hip.zd))

    def advect(self, u, v, w, u_mid, v_mid, w_mid, dt_t, extra_rhs=None):
        """semilag.py:584-671.  Returns (ustar, vstar, wstar) -- persistent
        ghost-shaped buffers with interiors set (ghosts not refreshed); with
        extra_rhs, a 4th element: list of interior-shaped triples interpolated
        at the same feet."""
        dt = float(dt_t)
        plan = self.plan
        hip = self._hip
        hip.nclamp.zero_()
        if extra_rhs is None:
            plan.advect(u, v, w, u_mid, v_mid, w_mid, dt, self.ustar, self.vstar, self.wstar,
                        hip.nclamp)
            self.n_clamped_last = hip.nclamp[0].clone()
            return self.ustar, self.vstar, self.wstar
        hip.set_mid(u_mid, v_mid, w_mid)
        fields = {"u": u, "v": v, "w": w}
        outs = {"u": self.ustar, "v": self.vstar, "w": self.wstar}
        extras = [dict() for _ in extra_rhs]
        nx, ny, nz = self.nx, self.ny, self.nz
        for ic, comp in enumerate("uvw"):
            # departure() counts the clamped feet once per component; the
            # extras are gathered at the same feet first and cloned, since
            # gather() reuses one output buffer (semilag.py:622-638)
            n = hip.departure(comp, dt)
            shape = plan.arrival_shape(comp)
            for t, triple in enumerate(extra_rhs):
                extras[t][comp] = hip.gather(comp, triple[ic], n).clone().reshape(shape)
            vals = hip.gather(comp, fields[comp], n)
            if comp == "w":
                outs[comp][1:nx + 1, 1:ny + 1, 1:nz] = vals
            else:
                outs[comp][1:nx + 1, 1:ny + 1, 1:nz + 1] = vals
        hip.set_mid(None, None, None)
        self.n_clamped_last = hip.nclamp[0].clone()
        return (self.ustar, self.vstar, self.wstar,
                [(e["u"], e["v"], e["w"]) for e in extras])
