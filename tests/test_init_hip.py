"""Device-side seeded initialisation (slc_init_solenoidal) against the numpy
synthesis of the same modes.

slchannel_amd.initflow.solenoidal_fields builds the seeded solenoidal
perturbation either with numpy FFTs on the host (the numpy path, which the
closed-form checks of the CPU suite pin: solenoidal modes, wall values, the
u'rms scale) or, given a plan, on the device by direct summation over the
wavenumber bins (include/slc.h).  Same modes, amplitudes and scale, summed in
another order: bar 1e-12 of the largest perturbation magnitude, ghosts
exactly zero, repeated calls bit-identical.  The reference has no seeded
initialiser (initflow.py:123-223), so there is nothing of its to compare.
"""

import math

import numpy as np
import pytest
import torch

from slchannel_amd import SLChannelFlow
from slchannel_amd.initflow import solenoidal_fields

pytestmark = pytest.mark.gpu

TOL = 1e-12

CASES = [
    # nx, ny, nz, stretching, spectrum, kmin, kmax, urms
    (24, 16, 20, "symmetric", "flat", 1.0, 8.0, 0.1),
    (32, 24, 32, "bottom", "k-5/3", 1.0, 6.0, 0.15),
    (20, 18, 16, "symmetric", "flat", 2.0, 4.0, 0.05),
    (33, 17, 24, "symmetric", "k-5/3", 1.0, 5.0, 0.12),
]


def _flow(nx, ny, nz, stretching):
    cfg = {"grid": {"nx": nx, "ny": ny, "nz": nz},
           "domain": {"Lx": 4 * math.pi, "Ly": 2 * math.pi, "Lz": 2.0,
                      "stretching_type": stretching},
           "flow": {"Re": 2800.0, "U_bulk": 1.0, "gamma": 1.5},
           "sl": {"interp_order": 6}}
    return SLChannelFlow(config=cfg, device="cuda:0")


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}x{c[1]}x{c[2]}-{c[3]}-{c[4]}")
def test_device_init_matches_numpy(check, case):
    nx, ny, nz, stretching, spectrum, kmin, kmax, urms = case
    f = _flow(nx, ny, nz, stretching)
    args = (nx, ny, nz, f.Lx, f.Ly, f.Lz, f.z_f.numpy(), f.z_c.numpy(), 4321)
    kw = dict(urms=urms, kmin=kmin, kmax=kmax, spectrum=spectrum, U_bulk=1.0)
    ref = solenoidal_fields(*args, **kw)
    dev = solenoidal_fields(*args, plan=f.plan, **kw)
    dev2 = solenoidal_fields(*args, plan=f.plan, **kw)
    torch.cuda.synchronize()
    zeta = 2.0 * f.z_c.numpy()[1:nz + 1] / f.Lz - 1.0
    base = 1.5 * (1.0 - zeta ** 2)
    for name, r, d, d2 in zip("uvw", ref, dev, dev2):
        d = d.cpu().numpy()
        pert = r.copy()
        if name == "u":
            pert[1:nx + 1, 1:ny + 1, 1:nz + 1] -= base[None, None, :]
        scale = np.abs(pert).max()
        err = np.abs(d - r).max() / scale
        check(f"{name}: device == numpy within {TOL:g} of max|perturbation|", err <= TOL,
              f"{err:.1e} (max|pert| {scale:.3g})")
        interior = np.zeros(r.shape, dtype=bool)
        interior[1:nx + 1, 1:ny + 1, 1:(nz if name == "w" else nz + 1)] = True
        check(f"{name}: ghosts exactly zero", np.all(d[~interior] == 0.0))
        check(f"{name}: repeated call bit-identical",
              np.array_equal(d, d2.cpu().numpy()))
    up = [(t.cpu().numpy() if torch.is_tensor(t) else t) for t in dev]
    up[0] = up[0].copy()
    up[0][1:nx + 1, 1:ny + 1, 1:nz + 1] -= base[None, None, :]
    ss = (np.sum(up[0] ** 2) + np.sum(up[1] ** 2) + np.sum(up[2] ** 2)) / (
        nx * ny * nz * 2 + nx * ny * (nz - 1))
    check("u'rms of the device perturbation", abs(math.sqrt(ss) - urms) <= 1e-12 * urms,
          f"{math.sqrt(ss):.15g} vs {urms}")


def test_initialize_is_device_built(check):
    """SLChannelFlow.initialize goes through slc_init_solenoidal: its projected
    fields equal those of the numpy fields projected by the same plan."""
    f = _flow(24, 16, 20, "symmetric")
    f.initialize(seed=77, urms=0.1, kmax=8.0)
    g = _flow(24, 16, 20, "symmetric")
    u, v, w = solenoidal_fields(24, 16, 20, g.Lx, g.Ly, g.Lz, g.z_f.numpy(), g.z_c.numpy(), 77,
                                urms=0.1, kmax=8.0)
    g.set_fields(u, v, w)
    g.apply_bc_uvw()
    g.project_initial()
    torch.cuda.synchronize()
    for name in "uvw":
        a, b = getattr(f, name), getattr(g, name)
        err = (a - b).abs().max().item() / b.abs().max().item()
        check(f"{name}: initialize == numpy fields projected", err <= TOL, f"{err:.1e}")
    div = f.divergence().abs().max().item()
    check("initialized field divergence-free", div < 1e-11, f"{div:.1e}")


def test_init_argument_errors(check):
    f = _flow(16, 12, 12, "symmetric")
    dev = f.device
    z = dict(dtype=torch.float64, device=dev)
    u = torch.zeros(f.shapes["u"], **z)
    v = torch.zeros(f.shapes["v"], **z)
    w = torch.zeros(f.shapes["w"], **z)
    ab = torch.zeros(5 * 9 * 9 * 2 * 2, **z)
    zt = torch.zeros(2 * 2 * 12 + 2 * 11 + 12, **z)
    ph = torch.zeros(3 * 9 * (16 + 12) * 2, **z)
    for W, L, nwork, want in ((8, 2, 10 ** 6, "odd"), (9, 2, 16, "work buffer")):
        work = torch.zeros(nwork, **z)
        try:
            f.plan.init_solenoidal(W, L, ab, zt, ph, 0.1, work, u, v, w)
            msg = ""
        except RuntimeError as e:
            msg = str(e)
        check(f"W={W}, work {nwork * 8} B rejected ({want})", want in msg, msg)
    need = f.plan.init_solenoidal_work(9)
    check("work size formula", need == 16 * 3 * 81 * 12 + 16 * 3 * 9 * 12 * 12 + 8 * 3 * 1024,
          need)
