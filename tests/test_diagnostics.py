"""Run-loop diagnostics (slc_diagnostics) and the blow-up guard.

CPU: the oracle's compute_u_tau / compute_cfl_fused / run_diagnostics
(utils.py:116-146, operators.py:839-940, solver.py:1355-1479) pinned by
closed forms: uniform and linear fields with known CFL, divergence and
friction velocity, the ghost rows inside the reference's u_tau plane mean,
NaN propagation, the non-finite count; BlowupGuard's three-strike rule
(solver.py:1202-1213).

GPU: the device diagnostics against the oracle on random and stepped
fields -- max|div| and dti bit-identical (same element operations, exact
maxima), u_tau to 1e-14, the non-finite count exact with NaN/Inf planted in
u, v, w and p.
"""

import math

import numpy as np
import pytest
import torch

import oracle.oracle as orc
from slchannel_amd.diagnostics import BlowupGuard


def _grid(nx=8, ny=6, nz=10, gamma=1.5, Lz=2.0):
    z_f, z_c, dz_f, dz_c = orc.generate_grid(gamma, nz, Lz)
    return dict(nx=nx, ny=ny, nz=nz, dx=4 * math.pi / nx, dy=2 * math.pi / ny,
                z_f=np.asarray(z_f), z_c=np.asarray(z_c), dz_f=np.asarray(dz_f),
                dz_c=np.asarray(dz_c))


def _zeros(g):
    nx, ny, nz = g["nx"], g["ny"], g["nz"]
    return (np.zeros((nx + 1, ny + 2, nz + 2)), np.zeros((nx + 2, ny + 1, nz + 2)),
            np.zeros((nx + 2, ny + 2, nz + 1)), np.zeros((nx + 2, ny + 2, nz + 2)))


def _cfl(g, u, v, w):
    return orc.compute_cfl_fused(u, v, w, g["nx"], g["ny"], g["nz"], g["dx"], g["dy"],
                                 g["dz_f"], g["dz_c"])


def test_cfl_closed_forms(check):
    g = _grid()
    u, v, w, _ = _zeros(g)
    u[:] = 0.7
    check("uniform u: dti = U/dx", abs(_cfl(g, u, v, w) - 0.7 / g["dx"]) < 1e-15 * 0.7 / g["dx"],
          f"{_cfl(g, u, v, w)!r}")
    v[:] = -0.3
    exp = 0.7 / g["dx"] + 0.3 / g["dy"]
    check("uniform u, v: dti = |U|/dx + |V|/dy", abs(_cfl(g, u, v, w) - exp) < 1e-14 * exp, "")
    u, v, w, _ = _zeros(g)
    nz = g["nz"]
    w[:, :, 1:nz] = 0.2  # interior faces; wall faces stay 0
    dzf, dzc = g["dz_f"], g["dz_c"]
    exp = max(0.2 / dzc[1:nz].min(), 0.2 / dzf[1:nz - 1].min(),
              0.1 / dzf[0], 0.1 / dzf[nz - 1])
    got = _cfl(g, u, v, w)
    check("uniform interior w: z faces vs half-weighted wall cells", abs(got - exp) < 1e-14 * exp,
          f"{got!r} vs {exp!r}")


def test_div_and_u_tau_closed_forms(check):
    g = _grid()
    nx, ny, nz = g["nx"], g["ny"], g["nz"]
    u, v, w, p = _zeros(g)
    a = 0.37
    u[:] = (a * g["dx"] * np.arange(nx + 1))[:, None, None]
    d = orc.run_diagnostics(u, v, w, p, nx, ny, nz, g["dx"], g["dy"], g["dz_f"], g["dz_c"],
                            g["z_c"], 1e-3, 0.01)
    check("linear u: max|div| = slope", abs(d[0] - a) < 1e-13, f"{d[0]!r}")
    # u = G z_c at every (x, y) -> u_tau_bot = sqrt(nu G), top scaled by z_c[nz]/z_c[1]
    G, nu = 0.8, 2e-3
    u, v, w, p = _zeros(g)
    u[:] = (G * g["z_c"])[None, None, :]
    bot = math.sqrt(nu * G)
    top = math.sqrt(nu * G * g["z_c"][nz] / g["z_c"][1])
    got = orc.compute_u_tau(u, g["z_c"], nu)
    check("u_tau both walls", abs(got - 0.5 * (bot + top)) < 1e-15, f"{got!r}")
    check("u_tau free-slip top: bottom only",
          abs(orc.compute_u_tau(u, g["z_c"], nu, "neumann") - bot) < 1e-15, "")
    # the reference's plane mean runs over the ghost rows too
    u[:, 0, :] = 0.0
    exp = math.sqrt(nu * G * (ny + 1) / (ny + 2))
    check("u_tau plane mean includes the y ghost rows",
          abs(orc.compute_u_tau(u, g["z_c"], nu, "neumann") - exp) < 1e-15, "")


def test_nonfinite_and_nan_propagation(check):
    g = _grid()
    nx, ny, nz = g["nx"], g["ny"], g["nz"]
    u, v, w, p = _zeros(g)
    u[3, 2, 4] = np.nan
    w[0, 0, 0] = np.inf      # a ghost entry counts too
    p[1, 1, 1] = -np.inf
    d = orc.run_diagnostics(u, v, w, p, nx, ny, nz, g["dx"], g["dy"], g["dz_f"], g["dz_c"],
                            g["z_c"], 1e-3, 0.01)
    check("non-finite entries counted", d[4] == 3.0, f"{d[4]}")
    check("NaN propagates to max|div| and the CFL", np.isnan(d[0]) and np.isnan(d[3]), f"{d}")


def test_blowup_guard_three_strikes(check):
    gd = BlowupGuard(0.05, factor=2.0)
    seq = [0.11, 0.12, 0.09, 0.11, 0.12, 0.13]
    out = [gd.check([0, x, 0, 0, 0]) for x in seq]
    check("single spikes reset, three in a row trip",
          out == [None, None, None, None, None, "blowup"], f"{out}")
    check("non-finite trips at once", BlowupGuard(0.05).check([0, 0.05, 0, 0, 1.0]) == "nonfinite",
          "")


@pytest.mark.gpu
def test_device_diagnostics_match_oracle(check):
    from slchannel_amd import HipPlan, SLChannelFlow

    cfg = {"grid": {"nx": 40, "ny": 36, "nz": 48},
           "domain": {"Lx": 4 * math.pi, "Ly": 2 * math.pi, "Lz": 2.0},
           "flow": {"Re": 3000.0, "U_bulk": 1.0, "gamma": 1.5},
           "sl": {"interp_order": 6}}
    f = SLChannelFlow(config=cfg)
    f.initialize(seed=11, urms=0.1, kmax=6.0)
    for _ in range(3):
        f.step_sl_bdf2(0.02)
    z_c, dz_f, dz_c = f.z_c.numpy(), f.dz_f.numpy(), f.dz_c.numpy()
    rng = np.random.default_rng(4)
    cases = [("stepped flow", [a.cpu().numpy() for a in (f.u, f.v, f.w, f.p)])]
    cases.append(("random", [rng.standard_normal(a.shape) for a in cases[0][1]]))
    bad = [a.copy() for a in cases[1][1]]
    bad[0][5, 3, 7] = np.nan
    bad[1][0, 0, 0] = np.inf
    bad[2][9, 9, 9] = -np.inf
    bad[3][2, 2, 2] = np.nan
    cases.append(("NaN/Inf planted", bad))
    for top in ("dirichlet", "neumann"):
        plan = f.plan if top == "dirichlet" else HipPlan(
            40, 36, 48, f.Lx, f.Ly, f.Lz, f.z_f, f.z_c, f.gamma, order=6,
            top_wall_bc_type="neumann", nu=f.nu, device=f.device)
        for name, (u, v, w, p) in cases:
            out = torch.empty(5, dtype=torch.float64, device="cuda")
            cu = [torch.tensor(a, device="cuda") for a in (u, v, w, p)]
            plan.diagnostics(*cu, 0.02, out)
            got = out.cpu().numpy()
            ref = orc.run_diagnostics(u, v, w, p, 40, 36, 48, f.dx, f.dy, dz_f, dz_c, z_c, f.nu,
                                      0.02, top)
            same = lambda a, b: (a == b) or (np.isnan(a) and np.isnan(b))  # noqa: E731
            check(f"{top} {name}: max|div| bit-identical", same(got[0], ref[0]),
                  f"{got[0]!r} vs {ref[0]!r}")
            check(f"{top} {name}: CFL bit-identical", same(got[2], ref[2]) and same(got[3], ref[3]),
                  f"{got[3]!r} vs {ref[3]!r}")
            check(f"{top} {name}: u_tau", abs(got[1] - ref[1]) <= 1e-14 * abs(ref[1]),
                  f"{got[1]!r} vs {ref[1]!r}")
            check(f"{top} {name}: non-finite count", got[4] == ref[4], f"{got[4]} vs {ref[4]}")
    row = f.diagnostics(dt=0.02)
    md = f.divergence().abs().max()
    check("SLChannelFlow.diagnostics max|div| == divergence()", row[0].item() == md.item(),
          f"{row[0].item()!r} vs {md.item()!r}")


@pytest.mark.gpu
def test_device_diagnostics_vs_reference_fixture(check, golden):
    """slc_diagnostics against the reference's compute_u_tau (both top-wall
    types) and compute_cfl_fused (tests/golden/diag_small.npz)."""
    from slchannel_amd import HipPlan

    g = golden("diag_small")
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    u, v, w = (torch.tensor(g[c], device="cuda") for c in "uvw")
    for top in ("dirichlet", "neumann"):
        plan = HipPlan(nx, ny, nz, float(g["Lx"]), float(g["Ly"]), float(g["Lz"]), g["z_f"],
                       g["z_c"], float(g["gamma"]), order=6, top_wall_bc_type=top,
                       nu=float(g["nu"]), device="cuda:0")
        out = torch.empty(5, dtype=torch.float64, device="cuda")
        plan.diagnostics(u, v, w, None, 0.1, out)
        got = out.tolist()
        ref = float(g[f"u_tau_{top}"])
        check(f"{top}: u_tau within 1e-15 of the reference", abs(got[1] - ref) <= 1e-15 * ref,
              f"{got[1]!r} vs {ref!r}")
        check(f"{top}: dti bit-identical to compute_cfl_fused", got[3] == float(g["dti"]),
              f"{got[3]!r} vs {float(g['dti'])!r}")
