"""The drop-in boundary at the reference's inner plugin protocol.

The reference's ``SLAdvector.advect`` (slchannel/semilag.py:604-666) drives its
fast path object (``TritonSL``, semilag_triton.py:347-455) like this:

  * fill the trajectory node buffers ``mbuf`` (semilag.py:604-606);
  * ``nclamp.zero_()``;
  * per component: fill ``fbuf`` (and the extras' node buffers), ``n =
    departure(comp, dt)``, each extra ``gather(comp, ebuf, n).clone()`` (the
    output buffer is reused), then ``vals = gather(comp, fbuf, n)`` written into
    the interior of the persistent output;
  * ``n_clamped = nclamp[0].long()``.

``RefProtocolAdvector`` below is a stand-in for the reference's advector with
exactly those attributes (packed node buffers, ``_fill``'s slices) and that
call sequence, with ``slchannel_amd.HipSL`` bound where ``TritonSL`` sits.
It is checked against the reference-written golden advect vectors, against
the oracle for the extras, and against this package's own ``SLAdvector``.

Bars: advected fields within 1e-13 of the max (the reference's eager fp64
path wrote the fixtures); clamp count equal; extras gathered at the same feet
bit-identical to the main gather of the same field; node-buffer departure and
gathers bit-identical to the ghost-shaped ones.
"""

import numpy as np
import pytest
import torch

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

CASES = ["o4", "o6", "o6_bottom", "o4_clamp", "o6_traj4"]


def cuda(a):
    return torch.as_tensor(np.ascontiguousarray(a)).to(torch.float64).cuda()


def rel_err(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = b.detach().cpu().numpy() if isinstance(b, torch.Tensor) else np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


class RefProtocolAdvector:
    """The reference advector's attributes and advect() call sequence
    (semilag.py:133-381, 387-394, 584-671), with HipSL as the fast path."""

    def __init__(self, g):
        from slchannel_amd import HipSL

        self.nx, self.ny, self.nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
        self.Lx, self.Ly, self.Lz = float(g["Lx"]), float(g["Ly"]), float(g["Lz"])
        self.dx, self.dy = self.Lx / self.nx, self.Ly / self.ny
        self.gamma = float(g["gamma"])
        self.stretching_type = str(g["stretching"])
        self.order = int(g["order"])
        self.traj_order = int(g["traj_order"]) if "traj_order" in g else 2
        self.n_traj_iters = 2
        self.top_wall_bc_type = "dirichlet"
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.z_f = torch.as_tensor(g["z_f"])
        self.z_c = torch.as_tensor(g["z_c"])
        self._tanh_g = torch.tanh(torch.tensor(self.gamma, dtype=torch.float64)).item()
        nx, ny, nz = self.nx, self.ny, self.nz
        node = {"u": (nx, ny, nz + 2), "v": (nx, ny, nz + 2), "w": (nx, ny, nz + 1)}
        f64 = dict(dtype=torch.float64, device=self.device)
        self.fbuf = {c: torch.empty(node[c], **f64) for c in "uvw"}
        self.mbuf = {c: torch.empty(node[c], **f64) for c in "uvw"}
        self._ebuf = []
        self.ustar = torch.zeros(nx + 1, ny + 2, nz + 2, **f64)
        self.vstar = torch.zeros(nx + 2, ny + 1, nz + 2, **f64)
        self.wstar = torch.zeros(nx + 2, ny + 2, nz + 1, **f64)
        self.arrival_shape = {"u": (nx, ny, nz), "v": (nx, ny, nz), "w": (nx, ny, nz - 1)}
        self._triton = HipSL(self)

    def _fill(self, buf, comp, field):
        nx, ny = self.nx, self.ny
        if comp == "u":
            buf.copy_(field[0:nx, 1:ny + 1, :])
        elif comp == "v":
            buf.copy_(field[1:nx + 1, 0:ny, :])
        else:
            buf.copy_(field[1:nx + 1, 1:ny + 1, :])

    def advect(self, u, v, w, u_mid, v_mid, w_mid, dt_t, extra_rhs=None):
        self._fill(self.mbuf["u"], "u", u_mid)
        self._fill(self.mbuf["v"], "v", v_mid)
        self._fill(self.mbuf["w"], "w", w_mid)
        fields = {"u": u, "v": v, "w": w}
        outs = {"u": self.ustar, "v": self.vstar, "w": self.wstar}
        extras = [dict() for _ in extra_rhs] if extra_rhs is not None else None
        while extra_rhs is not None and len(self._ebuf) < len(extra_rhs):
            self._ebuf.append({c: torch.empty_like(self.fbuf[c]) for c in "uvw"})
        self._triton.nclamp.zero_()
        dt_f = float(dt_t)
        for ic, comp in enumerate("uvw"):
            shape = self.arrival_shape[comp]
            self._fill(self.fbuf[comp], comp, fields[comp])
            if extras is not None:
                for t, triple in enumerate(extra_rhs):
                    self._fill(self._ebuf[t][comp], comp, triple[ic])
            n = self._triton.departure(comp, dt_f)
            if extras is not None:
                for t in range(len(extra_rhs)):
                    extras[t][comp] = self._triton.gather(comp, self._ebuf[t][comp],
                                                          n).clone().reshape(shape)
            vals = self._triton.gather(comp, self.fbuf[comp], n)
            nx, ny, nz = self.nx, self.ny, self.nz
            if comp == "w":
                outs[comp][1:nx + 1, 1:ny + 1, 1:nz] = vals.reshape(shape)
            else:
                outs[comp][1:nx + 1, 1:ny + 1, 1:nz + 1] = vals.reshape(shape)
        self.n_clamped_last = self._triton.nclamp[0].long()
        if extras is not None:
            return self.ustar, self.vstar, self.wstar, [(e["u"], e["v"], e["w"]) for e in extras]
        return self.ustar, self.vstar, self.wstar


def _oracle(g):
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    Lx, Ly = float(g["Lx"]), float(g["Ly"])
    sl = orc.SLOracle(nx, ny, nz, Lx / nx, Ly / ny, Lx, Ly, float(g["Lz"]), g["z_f"], g["z_c"],
                      float(g["gamma"]), stretching_type=str(g["stretching"]),
                      order=int(g["order"]),
                      traj_order=int(g["traj_order"]) if "traj_order" in g else 2)
    return sl


def _interior(comp, a, nx, ny, nz):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else a
    return a[1:nx + 1, 1:ny + 1, 1:nz] if comp == "w" else a[1:nx + 1, 1:ny + 1, 1:nz + 1]


@pytest.mark.parametrize("name", CASES)
def test_reference_protocol_vs_golden(check, golden, name):
    """HipSL driven through the reference's inner call sequence on node
    buffers reproduces the reference's own advect output and clamp count."""
    g = golden("advect_" + name)
    adv = RefProtocolAdvector(g)
    u, v, w = cuda(g["u"]), cuda(g["v"]), cuda(g["w"])
    um, vm, wm = cuda(g["um"]), cuda(g["vm"]), cuda(g["wm"])
    dt = float(g["dt"])
    us, vs, ws = adv.advect(u, v, w, um, vm, wm, dt)
    for nm, a, b in (("u", us, g["ustar"]), ("v", vs, g["vstar"]), ("w", ws, g["wstar"])):
        e = rel_err(a, b)
        check(f"{name}: advect {nm} vs reference", e < 1e-13, f"rel={e:.1e}")
    ncl = int(adv.n_clamped_last.item())
    check(f"{name}: n_clamped", ncl == int(g["n_clamped"]), f"{ncl} vs {int(g['n_clamped'])}")
    # node-buffer feet == ghost-shaped feet of the same velocity (both clamped)
    hip = adv._triton
    plan = hip.plan
    for comp in "uvw":
        n = plan.n_arrival(comp)
        hip.nclamp.zero_()
        hip.departure(comp, dt)
        a = [t[:n].clone() for t in (hip.xd, hip.yd, hip.zd)]
        c_nodes = int(hip.nclamp.item())
        hip.set_mid(um, vm, wm)
        hip.nclamp.zero_()
        hip.departure(comp, dt)
        b = [t[:n].clone() for t in (hip.xd, hip.yd, hip.zd)]
        c_ghost = int(hip.nclamp.item())
        hip.set_mid(None, None, None)
        same = all(torch.equal(x, y) for x, y in zip(a, b))
        check(f"{name}: {comp} node-buffer feet == ghost-shaped feet", same and c_nodes == c_ghost,
              f"clamped {c_nodes} / {c_ghost}")
        zc = b[2].cpu().numpy()
        check(f"{name}: {comp} departure z clamped", bool(((zc >= plan.z_lo) &
                                                             (zc <= plan.z_hi)).all()), "")


@pytest.mark.parametrize("name", ["o4", "o6", "o4_clamp"])
def test_reference_protocol_extra_rhs(check, golden, name):
    """extra_rhs through the reference protocol: each extra is gathered at the
    same feet as the field, cloned before the next gather."""
    g = golden("advect_" + name)
    adv = RefProtocolAdvector(g)
    u, v, w = cuda(g["u"]), cuda(g["v"]), cuda(g["w"])
    um, vm, wm = cuda(g["um"]), cuda(g["vm"]), cuda(g["wm"])
    dt = float(g["dt"])
    nx, ny, nz = adv.nx, adv.ny, adv.nz
    # extra 0: the advected fields themselves; extra 1: the trajectory velocity
    us, vs, ws, extras = adv.advect(u, v, w, um, vm, wm, dt, extra_rhs=[(u, v, w), (um, vm, wm)])
    for ic, (comp, out) in enumerate(zip("uvw", (us, vs, ws))):
        e0 = extras[0][ic].cpu().numpy()
        check(f"{name}: extra (u,v,w) {comp} == main gather",
              np.array_equal(e0, _interior(comp, out, nx, ny, nz)), "bit-identical")
    ncl = int(adv.n_clamped_last.item())
    check(f"{name}: n_clamped counted once per component", ncl == int(g["n_clamped"]),
          f"{ncl} vs {int(g['n_clamped'])}")
    sl = _oracle(g)
    sl.set_tanh_g(adv._tanh_g)
    ref = sl.advect(g["um"], g["vm"], g["wm"], g["um"], g["vm"], g["wm"], dt)
    for ic, comp in enumerate("uvw"):
        e = rel_err(extras[1][ic], _interior(comp, ref[ic], nx, ny, nz))
        check(f"{name}: extra (um,vm,wm) {comp} vs oracle", e < 1e-13, f"rel={e:.1e}")


@pytest.mark.parametrize("name", ["o6", "o4_clamp"])
def test_sladvector_extra_rhs(check, golden, name):
    """slchannel_amd.SLAdvector.advect(extra_rhs=...) (semilag.py:584-671)."""
    from slchannel_amd import SLAdvector

    g = golden("advect_" + name)
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    Lx, Ly = float(g["Lx"]), float(g["Ly"])
    adv = SLAdvector(nx, ny, nz, Lx / nx, Ly / ny, Lx, Ly, float(g["Lz"]), g["z_f"], g["z_c"],
                     float(g["gamma"]), stretching_type=str(g["stretching"]),
                     order=int(g["order"]), device="cuda")
    u, v, w = cuda(g["u"]), cuda(g["v"]), cuda(g["w"])
    um, vm, wm = cuda(g["um"]), cuda(g["vm"]), cuda(g["wm"])
    dt = float(g["dt"])
    us, vs, ws, extras = adv.advect(u, v, w, um, vm, wm, dt, extra_rhs=[(um, vm, wm)])
    for nm, a, b in (("u", us, g["ustar"]), ("v", vs, g["vstar"]), ("w", ws, g["wstar"])):
        e = rel_err(a, b)
        check(f"{name}: advect {nm} vs reference", e < 1e-13, f"rel={e:.1e}")
    ncl = int(adv.n_clamped_last.item())
    check(f"{name}: n_clamped", ncl == int(g["n_clamped"]), f"{ncl} vs {int(g['n_clamped'])}")
    sl = _oracle(g)
    sl.set_tanh_g(adv._tanh_g)
    ref = sl.advect(g["um"], g["vm"], g["wm"], g["um"], g["vm"], g["wm"], dt)
    for ic, comp in enumerate("uvw"):
        got = extras[0][ic]
        check(f"{name}: extra {comp} arrival-shaped", tuple(got.shape) ==
              adv.plan.arrival_shape(comp), str(tuple(got.shape)))
        e = rel_err(got, _interior(comp, ref[ic], nx, ny, nz))
        check(f"{name}: extra {comp} vs oracle", e < 1e-13, f"rel={e:.1e}")
    # without extras: the fused slc_advect gives the same doubles
    us2, vs2, ws2 = (t.clone() for t in adv.advect(u, v, w, um, vm, wm, dt))
    for nm, a, b in (("u", us, us2), ("v", vs, vs2), ("w", ws, ws2)):
        check(f"{name}: {nm} with extras == fused advect", torch.equal(a, b), "bit-identical")


def test_hipsl_gather_node_buffer_layout(check, golden):
    """HipSL.gather on an (nx, ny, NZ) node buffer (slc_gather layout 1) equals
    the gather of the ghost-shaped field at the same feet."""
    g = golden("advect_o6")
    adv = RefProtocolAdvector(g)
    hip = adv._triton
    fields = {"u": cuda(g["u"]), "v": cuda(g["v"]), "w": cuda(g["w"])}
    hip.set_mid(cuda(g["um"]), cuda(g["vm"]), cuda(g["wm"]))
    for comp in "uvw":
        n = hip.departure(comp, float(g["dt"]))
        ghost = hip.gather(comp, fields[comp], n).clone()
        adv._fill(adv.fbuf[comp], comp, fields[comp])
        packed = hip.gather(comp, adv.fbuf[comp], n).clone()
        e = rel_err(packed, ghost)
        check(f"gather {comp}: node buffer vs ghost-shaped", e < 1e-15, f"rel={e:.1e}")
    with pytest.raises(ValueError):
        hip.gather("u", fields["u"], 7)
