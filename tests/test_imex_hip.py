"""Eulerian IMEX path (slc_imex_rhs, slc_ab2_update, SLChannelFlow.step_imex)
against the oracle (oracle.compute_momentum_rhs_fused_imex and
ChannelOracle.step_imex, pinned in test_imex_oracle.py).

* The explicit rhs bit-identical to the oracle (same element operations) on
  random fields of a ragged grid.
* Full IMEX steps (forward Euler, then AB2; Crank-Nicolson z-solves; FFT
  projection with dt; bulk shift), both top-wall types: fields within 1e-12
  of the oracle over 6 steps, max|div| < 1e-11, u_bulk = 1 after each step.
"""

import math

import numpy as np
import pytest
import torch

import oracle.oracle as orc

pytestmark = pytest.mark.gpu


def _cfg(nx, ny, nz, top="dirichlet", Re=2000.0):
    return {"grid": {"nx": nx, "ny": ny, "nz": nz},
            "domain": {"Lx": 4 * math.pi, "Ly": 2 * math.pi, "Lz": 2.0},
            "flow": {"Re": Re, "U_bulk": 1.0, "gamma": 1.5},
            "boundary_conditions": {"top_wall": {"type": top}},
            "sl": {"interp_order": 4}}


def test_rhs_bit_identical(check):
    from slchannel_amd import SLChannelFlow

    f = SLChannelFlow(config=_cfg(24, 20, 18))
    rng = np.random.default_rng(8)
    u, v, w = (rng.standard_normal(f.shapes[c]) for c in "uvw")
    f.set_fields(u, v, w)
    ru, rv, rw = f.compute_momentum_rhs_explicit_imex()
    ou, ov, ow = orc.compute_momentum_rhs_fused_imex(u, v, w, 24, 20, 18, f.dx, f.dy,
                                                     f.dz_c.numpy(), f.dz_f.numpy(), f.nu)
    for name, a, b in (("u", ru, ou), ("v", rv, ov), ("w", rw, ow)):
        a = a.cpu().numpy()
        check(f"rhs_{name} bit-identical to the oracle", np.array_equal(a, b),
              f"max|d|={np.abs(a - b).max():.1e} of {np.abs(b).max():.1e}")


@pytest.mark.parametrize("top", ["dirichlet", "neumann"])
def test_imex_steps_match_oracle(check, top):
    from slchannel_amd import SLChannelFlow
    from slchannel_amd.diagnostics import cfl_dt

    f = SLChannelFlow(config=_cfg(32, 24, 28, top=top))
    f.initialize(seed=77, urms=0.1, kmax=6.0)
    co = orc.ChannelOracle(32, 24, 28, f.Lx, f.Ly, f.Lz, f.gamma, 2000.0, order=4,
                           top_wall_bc_type=top, z_f=f.z_f.numpy(), z_c=f.z_c.numpy(),
                           tanh_g=f.plan.tanh_g)
    co.set_fields(f.u.cpu().numpy(), f.v.cpu().numpy(), f.w.cpu().numpy(), f.p.cpu().numpy())
    dt = cfl_dt(f, 0.3)
    worst, per = 0.0, []
    for n in range(6):
        ub, _ = f.step_imex(dt)
        co.step_imex(dt)
        e = max(float(np.abs(a.cpu().numpy() - b).max() / np.abs(b).max())
                for a, b in ((f.u, co.u), (f.v, co.v), (f.w, co.w), (f.p, co.p)))
        worst = max(worst, e)
        per.append(f"{e:.0e}")
        md = f.divergence().abs().max().item()
        check(f"{top} step {n + 1}: max|div|", md < 1e-11, f"{md:.2e}")
        check(f"{top} step {n + 1}: u_bulk", abs(ub.item() - 1.0) < 1e-14, f"{ub.item()!r}")
    check(f"{top}: u, v, w, p vs oracle over 6 steps", worst < 1e-12,
          f"max rel={worst:.1e} per step={per} dt={dt:.4f}")


@pytest.mark.parametrize("tag", ["sym", "bot"])
def test_rhs_vs_reference_fixture(check, golden, tag):
    """slc_imex_rhs against the reference's own compute_momentum_rhs_fused_imex
    outputs (tests/golden/imex_small.npz): bit-identical."""
    from slchannel_amd import HipPlan

    g = golden("imex_small")
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    top = str(g[f"{tag}_top"])
    plan = HipPlan(nx, ny, nz, float(g["Lx"]), float(g["Ly"]), float(g["Lz"]), g[f"{tag}_z_f"],
                   g[f"{tag}_z_c"], float(g["gamma"]),
                   stretching_type="symmetric" if tag == "sym" else "bottom", order=6,
                   top_wall_bc_type=top, nu=float(g[f"{tag}_nu"]), device="cuda:0")
    u, v, w = (torch.tensor(g[f"{tag}_{c}"], device="cuda") for c in "uvw")
    out = [torch.empty_like(t) for t in (u, v, w)]
    plan.imex_rhs(u, v, w, *out)
    for name, a in zip(("ru", "rv", "rw"), out):
        a, b = a.cpu().numpy(), g[f"{tag}_{name}"]
        check(f"{tag} {name} bit-identical to the reference", np.array_equal(a, b),
              f"max|d|={np.abs(a - b).max():.1e} of {np.abs(b).max():.1e}")


def test_imex_steps_vs_reference_fixture(check, golden):
    """Three SLChannelFlow.step_imex steps against the reference's
    (tests/golden/imex_small.npz): within 1e-12 (Thomas vs PCR, rocFFT vs
    pocketfft rounding)."""
    from slchannel_amd import SLChannelFlow

    g = golden("imex_small")
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    f = SLChannelFlow(config={"grid": {"nx": nx, "ny": ny, "nz": nz},
                              "domain": {"Lx": float(g["Lx"]), "Ly": float(g["Ly"]),
                                         "Lz": float(g["Lz"])},
                              "flow": {"Re": float(g["step_Re"]), "U_bulk": 1.0,
                                       "gamma": float(g["gamma"])},
                              "sl": {"interp_order": 6}})
    f.set_fields(g["step_u0"], g["step_v0"], g["step_w0"], g["step_p0"])
    dt = float(g["step_dt"])
    for n in range(int(g["step_n"])):
        ub, fo = f.step_imex(dt)
        check(f"step {n + 1} u_bulk", abs(ub.item() - g["step_u_bulk"][n]) < 1e-14,
              f"{ub.item()!r}")
        for nm in "uvwp":
            a, b = getattr(f, nm).cpu().numpy(), g[f"step_{nm}_{n + 1}"]
            e = float(np.abs(a - b).max() / np.abs(b).max())
            check(f"step {n + 1} {nm} within 1e-12 of the reference", e < 1e-12, f"rel={e:.1e}")
