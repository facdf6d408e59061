"""Full-step analytic checks on the oracle and on the HIP layer.

  Stokes decay   tests/test_stokes_decay.py:14-52   decay within 1%, v = w = 0 (< 1e-12)
  divergence     tests/test_sl_divergence.py:53-62  max|div| < 1e-10 over 25 steps,
                                                     clamp fraction < 1e-2
(bulk forcing disabled for the Stokes mode as the reference test does).
"""

import math

import numpy as np
import pytest

from oracle import oracle as orc

BACKENDS = ["oracle", pytest.param("hip", marks=pytest.mark.gpu)]


def _cfg(nx, ny, nz, Re, gamma, order=4):
    return {"grid": {"nx": nx, "ny": ny, "nz": nz},
            "domain": {"Lx": 2 * math.pi, "Ly": 2 * math.pi, "Lz": 2.0},
            "flow": {"Re": Re, "U_bulk": 1.0, "gamma": gamma},
            "sl": {"interp_order": order}}


class _Flow:
    """Oracle or HIP stepper behind one interface (numpy in, numpy out)."""

    def __init__(self, be, cfg):
        self.be = be
        g = cfg["grid"]
        if be == "oracle":
            self.f = orc.ChannelOracle(g["nx"], g["ny"], g["nz"], cfg["domain"]["Lx"],
                                       cfg["domain"]["Ly"], cfg["domain"]["Lz"],
                                       cfg["flow"]["gamma"], cfg["flow"]["Re"],
                                       order=cfg["sl"]["interp_order"])
        else:
            from slchannel_amd import SLChannelFlow

            self.f = SLChannelFlow(config=cfg)

    @property
    def z_c(self):
        z = self.f.z_c
        return z.numpy() if hasattr(z, "numpy") else z

    @property
    def z_f(self):
        z = self.f.z_f
        return z.numpy() if hasattr(z, "numpy") else z

    def set(self, u, v, w):
        self.f.set_fields(u, v, w, None)
        if self.be == "oracle":
            self.f.p = np.zeros((self.f.nx + 2, self.f.ny + 2, self.f.nz + 2))
        self.f.apply_bc_uvw()

    def fields(self):
        if self.be == "oracle":
            return self.f.u, self.f.v, self.f.w
        return tuple(t.cpu().numpy() for t in (self.f.u, self.f.v, self.f.w))

    def div(self):
        f = self.f
        if self.be == "oracle":
            return orc.compute_divergence(f.u, f.v, f.w, f.nx, f.ny, f.nz, f.dx, f.dy, f.dz_f)
        return f.divergence().cpu().numpy()

    def nclamp(self):
        if self.be == "oracle":
            return self.f.n_clamped_last
        return int(self.f.sl.n_clamped_last.item())


@pytest.mark.parametrize("be", BACKENDS)
def test_stokes_decay(check, be):
    fl = _Flow(be, _cfg(8, 8, 96, 100.0, 1.3))
    nx, ny, nz = 8, 8, 96
    zc = fl.z_c
    u = np.broadcast_to(np.sin(math.pi * zc).reshape(1, 1, -1), (nx + 1, ny + 2, nz + 2)).copy()
    fl.set(u, np.zeros((nx + 2, ny + 1, nz + 2)), np.zeros((nx + 2, ny + 2, nz + 1)))
    fl.f._apply_bulk_forcing = lambda dt: (0.0, 0.0)
    dt, T = 0.02, 1.0
    for _ in range(round(T / dt)):
        fl.f.step_sl_bdf2(dt)
    u, v, w = fl.fields()
    decay = math.exp(-(1.0 / 100.0) * math.pi**2 * T)
    exact = decay * np.sin(math.pi * zc[1:-1])
    got = u[4, 4, 1:-1]
    e = np.abs(got - exact).max() / np.abs(exact).max()
    check("Stokes decay rate", e < 0.01, f"rel_err={e:.2e}")
    vm, wm = np.abs(v).max(), np.abs(w).max()
    check("v, w remain zero", vm < 1e-12 and wm < 1e-12, f"max|v|={vm:.1e} max|w|={wm:.1e}")


@pytest.mark.parametrize("order", [4, 6])
@pytest.mark.parametrize("be", BACKENDS)
def test_sl_divergence(check, be, order):
    from slchannel_amd.initflow import solenoidal_fields

    nx, ny, nz = 16, 16, 32
    fl = _Flow(be, _cfg(nx, ny, nz, 1000.0, 1.5, order))
    u, v, w = solenoidal_fields(nx, ny, nz, 2 * math.pi, 2 * math.pi, 2.0, fl.z_f, fl.z_c,
                                seed=7, urms=0.08, kmax=4.0)
    fl.set(u, v, w)
    fl.f.project_initial()
    max_div, ncl = 0.0, 0
    for _ in range(25):
        fl.f.step_sl_bdf2(0.02)
        max_div = max(max_div, float(np.abs(fl.div()).max()))
        ncl = fl.nclamp()
    u, v, w = fl.fields()
    finite = all(np.isfinite(a).all() for a in (u, v, w))
    frac = ncl / (3 * nx * ny * nz)
    check(f"SL bdf2 step order={order}", finite and max_div < 1e-10 and frac < 0.01,
          f"max|div|={max_div:.2e} clamp_frac={frac:.1e}")
