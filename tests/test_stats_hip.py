"""On-device turbulence statistics (include/slc_stats.h) against the oracle.

The oracle (oracle/turbstats.py) restates TurbulenceStats.accumulate_ and
finalize_statistics (turbstats.py:156-467) in numpy and is pinned by closed
forms in test_stats_oracle.py.  Bar: every running sum and every finalized
profile within 1e-12 of the oracle, relative to the largest magnitude among
the quantities of its order (first moments: the plane means; second
moments: the stresses and squared means; each spectrum: itself).  A
quantity that is zero up to rounding (the plane mean of w in a channel,
pinned by continuity) is thus measured on the scale of its peers, not on its
own.  The HIP plane reductions and rocFFT sum in another order than numpy.

Cases: seeded random fields on grids that exercise the chunking (one chunk;
several x, y and z chunks with ragged ends; odd nx and ny), extra spectra
planes, the reference's constructed plane-mean case through the product
class, a stepped channel flow, save/load of the state, determinism.
"""

import math

import numpy as np
import pytest
import torch

from oracle.turbstats import StatsOracle

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _fields(nx, ny, nz, seed, mean=(1.0, 0.05, 0.0)):
    g = torch.Generator().manual_seed(seed)
    u = mean[0] + 0.1 * torch.randn(nx + 1, ny + 2, nz + 2, generator=g, dtype=torch.float64)
    v = mean[1] + 0.1 * torch.randn(nx + 2, ny + 1, nz + 2, generator=g, dtype=torch.float64)
    w = mean[2] + 0.1 * torch.randn(nx + 2, ny + 2, nz + 1, generator=g, dtype=torch.float64)
    return u, v, w


def _stats(nx, ny, nz, planes=None, z_plus=15.0, Re_tau=180.0, top="dirichlet"):
    from slchannel_amd.grid import generate_grid
    from slchannel_amd.turbstats import TurbulenceStats

    Lx, Ly, Lz = 4 * math.pi, 2 * math.pi, 2.0
    z_f, z_c, dz_f, dz_c = generate_grid(1.5, nz, Lz, stretching_type="symmetric")
    return TurbulenceStats(nx, ny, nz, Lx, Ly, Lz, z_c, z_f, dz_c, dz_f, Lx / nx, Ly / ny,
                           1.0 / 3000.0, Re_tau, z_plus_target=z_plus, device="cuda",
                           top_wall_bc_type=top, spectra_z_planes=planes), float(z_c[1])


def _rel(a, b, scale=0.0):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), scale, 1e-300))


FIRST = (0, 5, 6)  # U, Vm, Wm sums; the other eight profile sums are second moments


def _compare(check, tag, ts, orc):
    names = ("U_sum", "uu_sum", "vv_sum", "ww_sum", "uw_sum", "Vm_sum", "Wm_sum", "Um2_sum",
             "Vm2_sum", "Wm2_sum", "UmWm_sum")
    s1 = max(np.max(np.abs(orc.prof[q])) for q in FIRST)
    s2 = max(np.max(np.abs(orc.prof[q])) for q in range(11) if q not in FIRST)
    worst = max((_rel(getattr(ts, n).cpu().numpy(), orc.prof[q], s1 if q in FIRST else s2), n)
                for q, n in enumerate(names))
    check(f"{tag}: profile sums within {TOL:g}", worst[0] < TOL, f"worst {worst[1]} {worst[0]:.1e}")
    spec = ts._spec.cpu().numpy()
    ws = max((_rel(spec[p, c], orc.spec[p, c]), p, c)
             for p in range(orc.spec.shape[0]) for c in range(4))
    check(f"{tag}: spectra sums within {TOL:g}", ws[0] < TOL,
          f"worst plane {ws[1]} comp {ws[2]} {ws[0]:.1e}")


def _compare_final(check, tag, got, ref):
    s2 = max(np.max(np.abs(ref[k])) for k in ("uu_mean", "vv_mean", "ww_mean", "uw_mean"))
    for k in ("U_mean", "uu_mean", "vv_mean", "ww_mean", "uw_mean"):
        r = _rel(got[k], ref[k], 0.0 if k == "U_mean" else s2)
        check(f"{tag}: finalized {k} within {TOL:g}", r < TOL, f"{r:.1e}")


@pytest.mark.parametrize("shape", [(16, 12, 24), (70, 40, 75), (33, 17, 9)])
def test_random_fields_match_oracle(check, shape):
    nx, ny, nz = shape
    ts, zc1 = _stats(nx, ny, nz, planes=[5.0, 40.0])
    planes = [(ts.k_bot, ts.k_top)] + [(p["k_bot"], p["k_top"]) for p in ts.extra_planes]
    orc = StatsOracle(nx, ny, nz, planes=planes)
    for s in range(3):
        u, v, w = _fields(nx, ny, nz, 100 * s + nx)
        orc.accumulate(u.numpy(), v.numpy(), w.numpy())
        ts.accumulate_statistics(u.cuda(), v.cuda(), w.cuda(), 0.05)
    _compare(check, f"{shape}", ts, orc)
    got = ts.finalize_statistics()
    ref = orc.finalize(ts.nu, zc1)
    _compare_final(check, f"{shape}", got, ref)
    check(f"{shape}: u_tau", abs(got["u_tau"] - ref["u_tau"]) <= TOL * abs(ref["u_tau"]),
          f"{got['u_tau']!r} vs {ref['u_tau']!r}")
    r = _rel(got["E_uu_2d"], ref["spectra"][0, 0])
    check(f"{shape}: finalized E_uu_2d", r < TOL, f"{r:.1e}")
    zp = ts.extra_planes[1]["zp"]
    r = _rel(got[f"E_uw_2d_z{zp:.0f}"], ref["spectra"][2, 3])
    check(f"{shape}: finalized extra-plane E_uw", r < TOL, f"{r:.1e}")


def test_reference_plane_mean_case(check):
    """tests/test_stats_plane_mean.py of the reference, through the GPU path."""
    nx, ny, nz = 16, 12, 24
    ts, _ = _stats(nx, ny, nz)
    i = torch.arange(nx + 1, dtype=torch.float64)
    face = torch.cos(2 * np.pi * i / nx)
    pattern_var = float((face[1:] ** 2).mean())
    cc = 0.5 * (face[:-1] + face[1:])
    pattern_var_filtered = float((cc ** 2).mean())
    u = torch.zeros(nx + 1, ny + 2, nz + 2, dtype=torch.float64, device="cuda")
    v = torch.zeros(nx + 2, ny + 1, nz + 2, dtype=torch.float64, device="cuda")
    w = torch.zeros(nx + 2, ny + 2, nz + 1, dtype=torch.float64, device="cuda")
    a = 0.37
    for n in range(8):
        u[:] = (face + (a if n % 2 == 0 else -a)).view(-1, 1, 1).cuda()
        ts.accumulate_statistics(u, v, w, 0.05)
    raw = float((ts.uu_sum / ts.n_samples)[nz // 2])
    st = ts.finalize_statistics()
    fixed = float(st["uu_mean"][nz // 2])
    check("corrected u'u' = pattern_var + a^2", abs(fixed - (pattern_var + a * a)) < 1e-12,
          f"{fixed!r}")
    check("raw accumulator biased", abs(raw - pattern_var) < 1e-12
          and abs(raw - (pattern_var + a * a)) > 0.9 * a * a, f"{raw!r}")
    check("u'u' at nodes", abs(raw - pattern_var_filtered) > 0.5 * (pattern_var - pattern_var_filtered),
          f"{raw!r} vs {pattern_var_filtered!r}")
    check("<u'w'> zero", abs(float(st["uw_mean"][nz // 2])) < 1e-14, f"{st['uw_mean'][nz // 2]:.1e}")
    check("w'w' zero", abs(float(st["ww_mean"][nz // 2])) < 1e-14, f"{st['ww_mean'][nz // 2]:.1e}")


def test_channel_flow_snapshots(check):
    """Statistics of a stepped channel (SLChannelFlow fields) vs the oracle."""
    from slchannel_amd import SLChannelFlow

    cfg = {"grid": {"nx": 48, "ny": 32, "nz": 40},
           "domain": {"Lx": 4 * math.pi, "Ly": 2 * math.pi, "Lz": 2.0},
           "flow": {"Re": 3000.0, "U_bulk": 1.0, "gamma": 1.5},
           "sl": {"interp_order": 6}}
    f = SLChannelFlow(config=cfg)
    f.initialize(seed=3, urms=0.1, kmax=6.0)
    ts, zc1 = _stats(48, 32, 40, top=f.top_wall_bc_type)
    orc = StatsOracle(48, 32, 40, planes=[(ts.k_bot, ts.k_top)])
    for _ in range(4):
        for _ in range(2):
            f.step_sl_bdf2(0.02)
        ts.accumulate_statistics(f.u, f.v, f.w, 0.05)
        orc.accumulate(f.u.cpu().numpy(), f.v.cpu().numpy(), f.w.cpu().numpy())
    _compare(check, "channel", ts, orc)
    got, ref = ts.finalize_statistics(), orc.finalize(ts.nu, zc1, f.top_wall_bc_type)
    _compare_final(check, "channel", got, ref)


def test_state_round_trip_and_determinism(check, tmp_path):
    nx, ny, nz = 24, 20, 30
    snaps = [_fields(nx, ny, nz, 7 + s) for s in range(5)]
    one, _ = _stats(nx, ny, nz, planes=[30.0])
    for u, v, w in snaps:
        one.accumulate_statistics(u.cuda(), v.cuda(), w.cuda())
    a, _ = _stats(nx, ny, nz, planes=[30.0])
    for u, v, w in snaps[:3]:
        a.accumulate_statistics(u.cuda(), v.cuda(), w.cuda())
    path = str(tmp_path / "state.npz")
    a.save_state(path)
    b, _ = _stats(nx, ny, nz, planes=[30.0])
    b.load_state(path)
    for u, v, w in snaps[3:]:
        b.accumulate_statistics(u.cuda(), v.cuda(), w.cuda())
    check("restart == uninterrupted (bit-identical sums)",
          torch.equal(one._prof, b._prof) and torch.equal(one._spec, b._spec),
          f"{(one._prof - b._prof).abs().max().item():.1e}")
    check("counters restored", b.n_samples == 5 and b.n_samples_extra == 5,
          f"{b.n_samples} {b.n_samples_extra}")
    c, _ = _stats(nx, ny, nz, planes=[30.0])
    for u, v, w in snaps:
        c.accumulate_statistics(u.cuda(), v.cuda(), w.cuda())
    check("deterministic (same input, same sums)",
          torch.equal(one._prof, c._prof) and torch.equal(one._spec, c._spec), "")
    bad, _ = _stats(nx + 2, ny, nz)
    try:
        bad.load_state(path)
        ok = False
    except ValueError:
        ok = True
    check("grid mismatch rejected", ok, "")


def test_reference_fixture(check, golden):
    """TurbulenceStats on the GPU against the reference's own accumulate and
    finalize outputs (tests/golden/stats_small.npz: three snapshots, the
    primary and two extra spectra planes), to the same bar as the oracle."""
    g = golden("stats_small")
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    ts, zc1 = _stats(nx, ny, nz, planes=[5.0, 40.0])
    check("plane indices == reference", (ts.k_bot, ts.k_top) == (int(g["k_bot"]), int(g["k_top"]))
          and [[p["k_bot"], p["k_top"]] for p in ts.extra_planes] == g["extra_k"].tolist(),
          f"{ts.k_bot, ts.k_top}")
    for n in range(3):
        ts.accumulate_statistics(*(torch.tensor(g[f"{c}{n}"], device="cuda") for c in "uvw"),
                                 0.05)
    names = ("U_sum", "uu_sum", "vv_sum", "ww_sum", "uw_sum", "Vm_sum", "Wm_sum", "Um2_sum",
             "Vm2_sum", "Wm2_sum", "UmWm_sum")
    s1 = max(np.max(np.abs(g[names[q]])) for q in FIRST)
    s2 = max(np.max(np.abs(g[n])) for q, n in enumerate(names) if q not in FIRST)
    worst = max((_rel(getattr(ts, n).cpu().numpy(), g[n], s1 if q in FIRST else s2), n)
                for q, n in enumerate(names))
    check(f"profile sums within {TOL:g} of the reference", worst[0] < TOL,
          f"worst {worst[1]} {worst[0]:.1e}")
    got = ts.finalize_statistics()
    ref = {k: g[f"fin_{k}"] for k in ("U_mean", "uu_mean", "vv_mean", "ww_mean", "uw_mean")}
    _compare_final(check, "reference", got, ref)
    check("u_tau", abs(got["u_tau"] - float(g["fin_u_tau"])) <= TOL * float(g["fin_u_tau"]),
          f"{got['u_tau']!r} vs {float(g['fin_u_tau'])!r}")
    for comp in ("uu", "vv", "ww", "uw"):
        r = _rel(got[f"E_{comp}_2d"], g[f"fin_E_{comp}_2d"])
        check(f"finalized E_{comp}_2d within {TOL:g}", r < TOL, f"{r:.1e}")
        for zp in (5, 40):
            key = f"E_{comp}_2d_z{zp}"
            r = _rel(got[key], g[f"fin_{key}"])
            check(f"finalized {key} within {TOL:g}", r < TOL, f"{r:.1e}")
