"""Serving loop (slchannel_amd/serve.py): continuous batching by HBM footprint
and device compute.

Requests of different grid sizes and step counts are served by one batched
ensemble under an HBM budget that admits only part of them at once, so runs
join while others are mid-flight (their BDF1 bootstrap inside the others'
BDF2 launches) and leave when done.  Runs that each fill the device go one
after the other; small ones ride along.  Each result must equal the same
request run alone on the plain per-run path (SLChannelFlow.step_sl_bdf2) bit
for bit: fields, pressure, bulk velocity, forcing and clamp count.
"""

import math

import pytest
import torch

pytestmark = pytest.mark.gpu


def _req(rid, nx, ny, nz, Re, steps, arrival=0.0):
    from slchannel_amd.serve import Request

    cfg = {"grid": {"nx": nx, "ny": ny, "nz": nz},
           "domain": {"Lx": 4 * math.pi, "Ly": 2 * math.pi, "Lz": 2.0},
           "flow": {"Re": Re, "U_bulk": 1.0, "gamma": 1.5},
           "sl": {"interp_order": 6}}
    return Request(rid=rid, config=cfg, n_steps=steps, seed=7000 + rid, arrival=arrival,
                   kmax=4.0)


def test_results_same_alone_and_in_the_mix(check):
    from slchannel_amd.ensemble import member_bytes
    from slchannel_amd.serve import Server, run_alone

    reqs = [_req(0, 24, 16, 24, 2800.0, 7), _req(1, 32, 24, 16, 4000.0, 4),
            _req(2, 24, 16, 24, 3500.0, 6, 0.01), _req(3, 16, 16, 32, 3000.0, 5, 0.02),
            _req(4, 32, 24, 16, 5000.0, 3, 0.03)]
    sizes = [member_bytes(r.config) for r in reqs]
    srv = Server(hbm_budget=1 << 40, keep_fields=True)
    srv.warmup([r.config for r in reqs])
    # room for two or three at once beside the server's fixed footprint (the
    # per-grid initialisers): later requests wait for a finished run
    fixed = srv.footprint()
    srv.budget = fixed + int(1.5 * (sorted(sizes)[-1] + sorted(sizes)[-2]))
    active_max = [0]
    foot_max = [0]
    log = []

    def note(msg):
        log.append(msg)
        active_max[0] = max(active_max[0], int(msg.split("active=")[1].split()[0]))
        foot_max[0] = max(foot_max[0], int(msg.split("footprint=")[1].split()[0]))

    results, metrics = srv.run(reqs, log=note)
    check("every request served", sorted(results) == [0, 1, 2, 3, 4] and
          all(r.status == "ok" for r in results.values()),
          f"{[(k, r.status, r.diag) for k, r in sorted(results.items())]}")
    check("admission bounded by the HBM budget", 1 < active_max[0] < len(reqs),
          f"max concurrent={active_max[0]} log={log}")
    check("footprint incl. pooled arenas and initialisers <= budget",
          fixed < foot_max[0] <= srv.budget, f"max {foot_max[0]} budget {srv.budget}")
    joined_mid = any(r.admitted > 0 and r.admitted < max(x.finished for x in results.values())
                     for r in results.values())
    check("runs joined while others were mid-flight", joined_mid, "")
    for r in reqs:
        alone, _ = run_alone(r)
        got = results[r.rid]
        same = all(torch.equal(got.fields[k], alone.fields[k]) for k in "uvwp")
        dev = max(((got.fields[k] - alone.fields[k]).abs().max().item()) for k in "uvwp")
        check(f"request {r.rid}: mix == alone ({r.n_steps} steps)", same and
              got.steps == alone.steps, f"max|d|={dev:.1e}")
        check(f"request {r.rid}: u_bulk, forcing, clamps == alone",
              got.stats[:3] == alone.stats[:3], f"{got.stats[:3]} vs {alone.stats[:3]}")
    check("latency percentiles", metrics["p99_step_ms"] >= metrics["p50_step_ms"] > 0,
          f"{metrics}")


def test_guards_stop_bad_runs_and_spare_the_rest(check):
    """The run-loop guards inside the serving loop (solver.py:1394-1479):
    a request with a NaN in its given initial state is rejected without a
    step; a run whose u_tau stays above the blow-up factor is stopped after
    three consecutive checks and its slot released; a healthy run in the
    same batch finishes and equals its run alone."""
    from slchannel_amd.ensemble import initial_fields
    from slchannel_amd.serve import Server, run_alone

    good = _req(0, 24, 16, 24, 2800.0, 6)
    bad = _req(1, 24, 16, 24, 3000.0, 6)
    u, v, w, p = initial_fields(bad.config, bad.seed, kmax=4.0)
    u = u.clone()
    u[3, 4, 5] = float("nan")
    bad.fields, bad.dt = (u, v, w, p), 0.02
    hot = _req(2, 32, 24, 16, 4000.0, 400)
    hot.u_tau_nominal = 1e-6  # every healthy u_tau is "above 2x nominal"
    srv = Server(keep_fields=True, guard_every=1)
    results, _ = srv.run([good, bad, hot])
    check("NaN initial state rejected unstepped",
          results[1].status == "rejected_nonfinite" and results[1].steps == 0,
          f"{results[1].status} steps={results[1].steps}")
    r2 = results[2]
    check("blow-up guard stops the run early", r2.status == "stopped_blowup" and
          3 <= r2.steps < 400, f"{r2.status} steps={r2.steps}")
    r0 = results[0]
    alone, _ = run_alone(good)
    same = all(torch.equal(r0.fields[k], alone.fields[k]) for k in "uvwp")
    check("healthy run finishes ok and == alone", r0.status == "ok" and same,
          f"{r0.status} diag={r0.diag}")
    check("final diagnostics finite, divergence-free",
          r0.diag[4] == 0.0 and r0.diag[0] < 1e-11 and r0.diag[1] > 0, f"{r0.diag}")


def test_invalid_requests_rejected_alone(check):
    """A malformed request is rejected into its Result; the others run."""
    from slchannel_amd.serve import Server

    good = _req(0, 24, 16, 24, 2800.0, 3)
    no_dt = _req(1, 24, 16, 24, 2800.0, 3)
    no_dt.fields = tuple(torch.zeros(s, dtype=torch.float64) for s in
                         ((25, 18, 26), (26, 17, 26), (26, 18, 25)))
    bad_shape = _req(2, 24, 16, 24, 2800.0, 3)
    bad_shape.fields = (torch.zeros(3, 3, 3),) * 3
    bad_shape.dt = 0.01
    tricubic = _req(3, 24, 16, 24, 2800.0, 3)
    tricubic.config["sl"]["traj_interp_order"] = 4
    srv = Server()
    results, _ = srv.run([good, no_dt, bad_shape, tricubic])
    check("good request ok", results[0].status == "ok" and results[0].steps == 3,
          results[0].status)
    for rid in (1, 2, 3):
        r = results[rid]
        check(f"request {rid} rejected_invalid", r.status == "rejected_invalid" and r.steps == 0,
              f"{r.status} {r.diag}")


def test_compute_aware_admission(check):
    """Runs that fill the device go one after the other in arrival order, in
    the solo lane (the single-run solver); a small request that arrives
    behind a waiting large one rides in the batch beside the running large
    one instead of queueing.  Every result equals its run alone."""
    from slchannel_amd.serve import Server, run_alone

    big_a = _req(0, 48, 32, 32, 2800.0, 12, 0.0)
    big_b = _req(1, 48, 32, 32, 3200.0, 3, 0.001)
    small = _req(2, 24, 16, 24, 3000.0, 3, 0.002)
    # 48x32x32 = 49152 grid points: two of them exceed the batch budget; the
    # 9216-point request is a rider
    srv = Server(keep_fields=True, batch_points=60000, small_points=20000)
    srv.warmup([r.config for r in (big_a, small)])
    log = []
    results, _ = srv.run([big_a, big_b, small], log=log.append)
    order = [int(m.split()[1]) for m in log]
    active_at = {int(m.split()[1]): int(m.split("active=")[1].split()[0]) for m in log}
    check("all served", all(results[i].status == "ok" for i in range(3)),
          f"{[results[i].status for i in range(3)]}")
    check("admitted in the order large, small, large", order == [0, 2, 1], f"{log}")
    # when the second large run is admitted the first has left: at most the
    # rider and the newcomer are active
    check("second large run waits for the first", active_at[1] <= 2, f"{log}")
    lanes = {int(m.split()[1]): m.split("lane=")[1] for m in log}
    check("large runs in the solo lane, the small one in the batch",
          lanes == {0: "solo", 1: "solo", 2: "ens"}, f"{lanes}")
    for r in (big_a, big_b, small):
        alone, _ = run_alone(r)
        same = all(torch.equal(results[r.rid].fields[k], alone.fields[k]) for k in "uvwp")
        check(f"request {r.rid}: == alone", same, "")
