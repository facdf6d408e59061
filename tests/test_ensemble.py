"""Batched ensemble (include/slc_ens.h) against the single-run path.

Every member of an ensemble must follow exactly the trajectory of the single
run (SLChannelFlow.step_sl_bdf2, itself pinned to the oracle and the
reference's golden vectors in test_hip_parity.py) from the same state: same
fields, pressure, bulk velocity, forcing and clamp counts.  The batched
kernels run the single-run device code per member and rocFFT transforms
each plane independently, so the bar is bit-identity; the test reports the
largest difference if it is not met.

Cases: uneven members (three grid sizes, two in one size class, different
Re and dt); members joining and leaving between steps (continuous
batching: a new member bootstraps with BDF1 while the others take BDF2
steps in the same launches); FP32-storage gathers; every tile shape;
hipGraph replay checked against eager stepping after every replay.
"""

import math

import pytest
import torch

pytestmark = pytest.mark.gpu


def _cfg(nx, ny, nz, Re, order=6, gamma=1.5, fp32=False):
    return {"grid": {"nx": nx, "ny": ny, "nz": nz},
            "domain": {"Lx": 4 * math.pi, "Ly": 2 * math.pi, "Lz": 2.0},
            "flow": {"Re": Re, "U_bulk": 1.0, "gamma": gamma},
            "sl": {"interp_order": order,
                   "interp_dtype": "fp32_accum64" if fp32 else "fp64"}}


MEMBERS = [
    (_cfg(24, 16, 24, 2800.0), 11, 0.03),
    (_cfg(32, 24, 16, 4000.0), 12, 0.025),
    (_cfg(24, 16, 24, 3500.0), 13, 0.02),
    (_cfg(16, 16, 32, 3000.0, gamma=1.8), 14, 0.03),
]


def _single(cfg, seed, fp32=False):
    from slchannel_amd import SLChannelFlow

    c = dict(cfg)
    c["sl"] = dict(cfg["sl"], interp_dtype="fp32_accum64" if fp32 else "fp64")
    f = SLChannelFlow(config=c)
    f.initialize(seed=seed, urms=0.1, kmax=4.0)
    return f


def _diff(a, b):
    return (a - b).abs().max().item() / max(b.abs().max().item(), 1e-300)


def _compare(check, tag, ens, mid, flow, ub=None):
    got = ens.get_fields(mid)
    same = all(torch.equal(got[k], getattr(flow, k)) for k in "uvwp")
    worst = max(_diff(got[k], getattr(flow, k)) for k in "uvwp")
    check(f"{tag}: fields == single run", same, f"max rel diff={worst:.1e}")
    st = ens.stats(mid).cpu()
    if ub is not None:
        check(f"{tag}: u_bulk == single run", st[0].item() == ub.item(),
              f"{st[0].item()!r} vs {ub.item()!r}")
    ncl = int(flow.sl.n_clamped_last.item())
    check(f"{tag}: clamp count == single run", int(st[2].item()) == ncl,
          f"{int(st[2].item())} vs {ncl}")


def _ensemble(storage=64, graph=False):
    from slchannel_amd.ensemble import Ensemble

    return Ensemble(order=6, storage=storage, graph=graph)


@pytest.mark.parametrize("storage", [64, 32])
def test_uneven_members_match_single_runs(check, storage):
    ens = _ensemble(storage)
    flows, ids = [], []
    for cfg, seed, dt in MEMBERS:
        f = _single(cfg, seed, fp32=storage == 32)
        ids.append(ens.add(cfg, dt, fields=(f.u, f.v, f.w, f.p)))
        flows.append(f)
    for n in range(5):
        ens.step()
        ubs = [f.step_sl_bdf2(dt)[0] for f, (_, _, dt) in zip(flows, MEMBERS)]
        for k, (mid, f) in enumerate(zip(ids, flows)):
            _compare(check, f"step {n + 1} member {k}", ens, mid, f, ubs[k])
    st = ens.stats(ids[0]).cpu()
    check("steps counted", int(st[3].item()) == 5, f"{st.tolist()}")


def test_long_and_short_columns_match_single_runs(check):
    """A member with 160 levels (4 unknowns per lane in the z-solve, one
    column per wavefront) beside short-column members (four columns per
    wavefront): the ensemble takes its general z-solve kernel, and every
    member stays bit-identical to its own run."""
    members = [(_cfg(16, 12, 160, 3000.0), 21, 0.02), (_cfg(24, 16, 24, 2800.0), 22, 0.03),
               (_cfg(20, 12, 100, 3500.0, gamma=1.8), 23, 0.025)]
    ens = _ensemble()
    flows, ids = [], []
    for cfg, seed, dt in members:
        f = _single(cfg, seed)
        ids.append(ens.add(cfg, dt, fields=(f.u, f.v, f.w, f.p)))
        flows.append(f)
    for n in range(3):
        ens.step()
        ubs = [f.step_sl_bdf2(dt)[0] for f, (_, _, dt) in zip(flows, members)]
        for k, (mid, f) in enumerate(zip(ids, flows)):
            _compare(check, f"step {n + 1} member {k}", ens, mid, f, ubs[k])


def test_join_and_leave_between_steps(check):
    ens = _ensemble()
    (ca, sa, dta), (cb, sb, dtb), (cc, sc, dtc) = MEMBERS[:3]
    fa, fb, fc = _single(ca, sa), _single(cb, sb), _single(cc, sc)
    ia = ens.add(ca, dta, fields=(fa.u, fa.v, fa.w, fa.p))
    ic = ens.add(cc, dtc, fields=(fc.u, fc.v, fc.w, fc.p))
    for _ in range(3):
        ens.step()
        fa.step_sl_bdf2(dta)
        fc.step_sl_bdf2(dtc)
    # b joins (BDF1 bootstrap inside BDF2 launches), c leaves
    ib = ens.add(cb, dtb, fields=(fb.u, fb.v, fb.w, fb.p))
    ens.remove(ic)
    for n in range(3):
        ens.step()
        ua = fa.step_sl_bdf2(dta)[0]
        ubb = fb.step_sl_bdf2(dtb)[0]
        _compare(check, f"after join {n + 1}: a", ens, ia, fa, ua)
        _compare(check, f"after join {n + 1}: b", ens, ib, fb, ubb)
    check("membership", len(ens) == 2, f"{len(ens)}")
    # d has c's geometry (other Re, dt, seed): it reuses c's released arena
    cd = _cfg(24, 16, 24, 5000.0)
    fd = _single(cd, 21)
    ida = ens.add(cd, 0.015, fields=(fd.u, fd.v, fd.w, fd.p))
    for n in range(3):
        ens.step()
        fa.step_sl_bdf2(dta)
        fb.step_sl_bdf2(dtb)
        ud = fd.step_sl_bdf2(0.015)[0]
        _compare(check, f"reused arena {n + 1}: d", ens, ida, fd, ud)
    _compare(check, "reused arena: a still exact", ens, ia, fa)


def test_every_tile_shape_and_autotune(check):
    from slchannel_amd.ensemble import OPT_TILE

    from slchannel_amd.ensemble import OPT_BDF2_FUSED

    results = {}
    for shape in (0, 1, 2, "split"):
        ens = _ensemble()
        if shape == "split":  # far pass + near pass instead of the fused pass
            ens.set_option(OPT_BDF2_FUSED, 0)
            check("split passes selected", ens.get_option(OPT_BDF2_FUSED) == 0, "")
        else:
            ens.set_option(OPT_TILE, shape)
        ids = []
        for cfg, seed, dt in MEMBERS[:2]:
            f = _single(cfg, seed)
            ids.append(ens.add(cfg, dt, fields=(f.u, f.v, f.w, f.p)))
        for _ in range(3):
            ens.step()
        results[shape] = [ens.get_fields(i) for i in ids]
        if shape == 2:
            rep, grids = ens.autotune(reps=2)
            check("autotune times every shape and grid mode",
                  sorted(rep) == [0, 1, 2] and sorted(grids) == [0, 4, 16], f"{rep} {grids}")
            # the report carries each float time exactly (%.9g), so the kept
            # shape's time is the minimum; equal times keep the first shape
            chosen = ens.get_option(OPT_TILE)
            check("autotune keeps the fastest", rep.get(chosen) == min(rep.values()),
                  f"times={rep} chosen={chosen}")
            ens.step()
            st = ens.stats(ids[0]).cpu()
            check("state unchanged by autotune (4 steps)", int(st[3].item()) == 4,
                  f"{st.tolist()}")
    for shape in (1, 2, "split"):
        same = all(torch.equal(a[k], b[k]) for a, b in zip(results[shape], results[0])
                   for k in "uvwp")
        check(f"tile shape {shape} == shape 0 (fused)", same, "")


def test_graph_replay_matches_eager(check):
    eager, graph = _ensemble(), _ensemble(graph=True)
    ids_e, ids_g = [], []
    for cfg, seed, dt in MEMBERS[:3]:
        f = _single(cfg, seed)
        ids_e.append(eager.add(cfg, dt, fields=(f.u, f.v, f.w, f.p)))
        ids_g.append(graph.add(cfg, dt, fields=(f.u, f.v, f.w, f.p)))
    for n in range(6):
        eager.step()
        graph.step()
        torch.cuda.synchronize()
        # unrelated eager work between replays
        torch.randn(1 << 16, device="cuda").sum().item()
        for k, (a, b) in enumerate(zip(ids_e, ids_g)):
            fa, fb = eager.get_fields(a), graph.get_fields(b)
            same = all(torch.equal(fa[x], fb[x]) for x in "uvwp")
            check(f"replay {n + 1} member {k}: graph == eager", same, "")
