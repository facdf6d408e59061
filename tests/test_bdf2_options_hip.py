"""The non-default step options on the HIP layer (solver.py:253-281, 846-1012;
semilag.py:179-209): sl.traj_interp_order 4 (tricubic trajectory samples, the
direct-load kernels), and the BDF2 options
sl.bdf2_xy_rhs 'lagged' (the predictor kernels take R^n) and
sl.bdf2_pressure 'inc' (p_ext = 2 p^n - p^{n-1} through slc_pressure_combine,
-dt_eff grad p_ext added to the predictor through slc_project before the
z-solves, p^{n+1} = p_ext + phi).  HIP against the oracle with the same
options every step, pressure included; the option checks run on the CPU.
"""

import math

import numpy as np
import pytest
import torch

from oracle import oracle as orc


def _cfg(order=4, **sl):
    return {"grid": {"nx": 32, "ny": 32, "nz": 24},
            "domain": {"Lx": 4 * math.pi, "Ly": 2 * math.pi, "Lz": 2.0},
            "flow": {"Re": 2800.0, "U_bulk": 1.0, "gamma": 1.5},
            "sl": {"interp_order": order, **sl}}


def _rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else a
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_solver_option_checks(check):
    from slchannel_amd import SLChannelFlow

    for sl, msg in (({"bdf2_pressure": "incremental"}, "bdf2_pressure"),
                    ({"bdf2_xy_rhs": "ab2"}, "bdf2_xy_rhs")):
        with pytest.raises(ValueError, match=msg):
            SLChannelFlow(config=_cfg(**sl))
    check("invalid option strings rejected before any device work", True, "")


@pytest.mark.gpu
@pytest.mark.parametrize("opts", [{"bdf2_xy_rhs": "lagged"}, {"bdf2_pressure": "inc"},
                                  {"bdf2_pressure": "inc", "bdf2_xy_rhs": "lagged"},
                                  {"traj_interp_order": 4}],
                         ids=["lagged", "inc", "inc+lagged", "traj4"])
def test_options_vs_oracle(check, opts):
    from slchannel_amd import SLChannelFlow
    from slchannel_amd.diagnostics import cfl_dt

    flow = SLChannelFlow(config=_cfg(**opts))
    flow.initialize(seed=1003, urms=0.1, kmax=6.0)
    co = orc.ChannelOracle(flow.nx, flow.ny, flow.nz, flow.Lx, flow.Ly, flow.Lz, flow.gamma,
                           2800.0, order=4, z_f=flow.z_f.numpy(), z_c=flow.z_c.numpy(),
                           tanh_g=flow.plan.tanh_g,
                           traj_order=opts.get("traj_interp_order", 2),
                           **{k: v for k, v in opts.items() if k.startswith("bdf2_")})
    co.set_fields(*(t.cpu().numpy() for t in (flow.u, flow.v, flow.w, flow.p)))
    dt = cfl_dt(flow, 3.0)
    worst_v, worst_p = 0.0, 0.0
    for n in range(8):
        ub, _ = flow.step_sl_bdf2(dt)
        co.step_sl_bdf2(dt)
        ev = max(_rel(a, b) for a, b in ((flow.u, co.u), (flow.v, co.v), (flow.w, co.w)))
        ep = _rel(flow.p, co.p)
        worst_v, worst_p = max(worst_v, ev), max(worst_p, ep)
        md = flow.divergence().abs().max().item()
        check(f"step {n + 1} max|div|", md < 1e-11, f"{md:.2e}")
        check(f"step {n + 1} u_bulk", abs(ub.item() - 1.0) < 1e-14, f"{ub.item()!r}")
    check("u, v, w vs oracle over 8 steps", worst_v < 1e-12, f"max rel={worst_v:.1e}")
    check("p vs oracle over 8 steps", worst_p < 1e-11, f"max rel={worst_p:.1e}")
    if opts.get("bdf2_pressure") == "inc":
        # the full pressure carries p_ext: it is not the last phi
        phi = flow._phi
        d = _rel(flow.p, phi.cpu().numpy())
        check("p = p_ext + phi differs from phi", d > 1e-6, f"rel {d:.1e}")


def test_ensemble_rejects_single_run_options(check):
    from slchannel_amd.ensemble import _config_grid

    for sl, msg in (({"bdf2_pressure": "inc"}, "noninc"), ({"bdf2_xy_rhs": "lagged"}, "extrap"),
                    ({"traj_interp_order": 4}, "trilinear")):
        with pytest.raises(ValueError, match=msg):
            _config_grid(_cfg(**sl))
    check("ensemble members keep the production options", True, "")
