"""Benchmark: SL-BDF2 steps/s of one channel at 256^3, triquintic, float64.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--order 6] [--n 256]

Workload (the single-GPU configuration of the headline metric): 256x256x256
channel, Lx = 4 pi, Ly = 2 pi, Lz = 2, tanh factor 1.5, Re_b = 2792.8,
interp_order 6, seeded Poiseuille + solenoidal k^-5/3 perturbation over
1 <= |k| <= 32 with u'rms = 0.15, dt+ = 0.2 (Re_tau = 180).  A "step" is one
full SL-BDF2 step (predictor, z-solves, FFT projection, bulk shift); the W
warm-up steps include the BDF1 bootstrap.  Inputs are resident in HBM.

Prints ONE JSON line.  Also reported:
  roofline      the dominant kernel -- the fused BDF2 pass of u (two-foot
                tracer, V^{n-1} gathered at the 2dt foot and V^n at the dt
                foot, xy right-hand side, BDF2 combination) -- timed with
                hipEvents on its launch stream over the timed region;
                algorithmic bytes per arrival point: U* 3x8 + V^{n-1} 8 +
                V^n 8 read, the prediction 8 written = 48 B (--bdf2-split:
                the far pass, 64 B); checked against the FP64 vector peak
                too (fp64_check);
  gather_GBps   the isolated order-P gather of u (slc_gather) at 40 B/point
                (3x8 B feet + 8 B field + 8 B out), feet from the 2dt foot;
                gather_GBps_by_order: the same field and feet at orders 4
                and 6 (a second plan of the other order);
  cpu_baseline  the CPU oracle (oracle/: numpy + OpenMP C) timed on this
                host for ONE BDF2 step of the same 256^3 state.
"""

import argparse
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

PEAK_HBM_GBPS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--order", type=int, default=6)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--nz", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 1 GiB device-to-device copy 3 times (a kernel of "
                         "known traffic, to scale the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--storage", type=int, default=64, choices=(64, 32),
                    help="bits of the gathered field copies: 64 (default) or 32 "
                         "(FP32-storage variant; adds an FP64 twin run and reports the "
                         "per-field difference)")
    ap.add_argument("--traj-order", type=int, default=2, choices=(2, 4),
                    help="trajectory sampling order (4: tricubic, the direct-load kernels)")
    ap.add_argument("--bdf2-split", action="store_true",
                    help="far pass, then near pass, instead of the fused BDF2 pass "
                         "(same doubles; for profiling the two passes)")
    ap.add_argument("--config", type=int, default=2, choices=(2, 3, 4, 5, 6, 7),
                    help="2: 256^3 channel (default, the headline workload); "
                         "3: 1024x1024x512 large box, CFL 4, plus the isolated gather "
                         "with feet displaced uniformly in +-5 cells; "
                         "4: 48-member uneven ensemble, batched vs one run at a time; "
                         "5: serving loop with Poisson arrivals (p50/p99 step latency); "
                         "6: on-device turbulence statistics; "
                         "7: the reference's advect microbenchmark (three-component "
                         "SLAdvector.advect at 768x768x180, orders 4 and 6)")
    ap.add_argument("--scheme", default="sl", choices=("sl", "imex"),
                    help="config 2: time the SL-BDF2 step (default) or the Eulerian IMEX "
                         "step at the reference's Eulerian CFL 0.28, with both schemes' "
                         "simulated time per second")
    ap.add_argument("--members", type=int, default=48, help="config 4: ensemble size")
    ap.add_argument("--ensemble-grids", default="workload", choices=("workload", "small"),
                    help="config 4: member grids -- the workload's four sizes, or small "
                         "launch-bound members (32x24x32, 48x32x48, 32x32x24)")
    ap.add_argument("--requests", type=int, default=24, help="config 5: number of requests")
    ap.add_argument("--rate", type=float, default=1.0, help="config 5: arrivals per second")
    ap.add_argument("--serve-grids", default="workload", choices=("workload", "small"),
                    help="config 5: request grids -- the workload's four sizes, or small "
                         "launch-bound ones (32x24x32, 48x32x48, 32x32x24)")
    ap.add_argument("--hbm-budget-gb", type=float, default=48.0,
                    help="config 5: HBM budget of the admission control")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); this library runs the "
              "channel on a single MI355X and has no multi-GPU path.", file=sys.stderr)
        sys.exit(2)
    import torch

    from slchannel_amd import SLChannelFlow

    if args.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.config == 3:
        return large_box(args)
    if args.config == 4:
        return ensemble_bench(args)
    if args.config == 5:
        return serve_bench(args)
    if args.config == 6:
        return stats_bench(args)
    if args.config == 7:
        return advect_bench(args)
    n, nz = args.n, args.nz or args.n
    Re_b, Re_tau, dtp = 2792.8, 180.0, 0.2
    cfg = {
        "grid": {"nx": n, "ny": n, "nz": nz},
        "domain": {"Lx": 4 * math.pi, "Ly": 2 * math.pi, "Lz": 2.0},
        "flow": {"Re": Re_b, "U_bulk": 1.0, "gamma": 1.5},
        "sl": {"interp_order": args.order,
               "interp_dtype": "fp32_accum64" if args.storage == 32 else "fp64",
               "traj_interp_order": args.traj_order},
    }
    flow = SLChannelFlow(config=cfg, device="cuda:0")
    flow.initialize(seed=2000, urms=0.15, kmin=1.0, kmax=32.0, spectrum="k-5/3")
    if args.bdf2_split:
        flow.plan.set_option(flow.plan.OPT_BDF2_FUSED, 0)
    nu = flow.nu
    u_tau = Re_tau * nu  # delta = 1
    dt = dtp * nu / (u_tau * u_tau)
    twin = None
    if args.storage == 32:
        cfg64 = dict(cfg, sl={"interp_order": args.order, "interp_dtype": "fp64"})
        twin = SLChannelFlow(config=cfg64, device="cuda:0")
        twin.set_fields(flow.u, flow.v, flow.w, flow.p)

    if args.scheme == "imex":
        return imex_bench(args, flow, dt)
    for _ in range(args.warmup):
        flow.step_sl_bdf2(dt)
    torch.cuda.synchronize()

    plan = flow.plan
    plan.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ub, fo = flow.step_sl_bdf2(dt)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = plan.profile_report()
    plan.profile(False)
    ms_step = (t1 - t0) * 1e3 / args.steps

    N_u = n * n * nz
    lab, kname, avg_ms, bytes_far = roofline_kernel(prof, N_u, args.storage)
    achieved = bytes_far / (avg_ms * 1e-3) / 1e9

    # isolated gather of u at the 2dt foot of the current state
    xd, yd, zd = (torch.empty(N_u, dtype=torch.float64, device="cuda") for _ in range(3))
    mids = [(2.0 * a - b).contiguous() for a, b in ((flow.u, flow.u_nm1), (flow.v, flow.v_nm1),
                                                    (flow.w, flow.w_nm1))]
    plan.departure("u", *mids, 2.0 * dt, xd, yd, zd)
    out = torch.empty(N_u, dtype=torch.float64, device="cuda")
    plan.gather("u", flow.u_nm1, xd, yd, zd, out)
    torch.cuda.synchronize()
    plan.profile(True)
    for _ in range(5):
        plan.gather("u", flow.u_nm1, xd, yd, zd, out)
    gp = plan.profile_report()
    plan.profile(False)
    g_cnt, g_tot = gp["gather_u"]
    gather_gbps = 40.0 * N_u / (g_tot / g_cnt * 1e-3) / 1e9
    gather_by_order = {str(args.order): round(gather_gbps, 1)}
    # the same gather (same field, same feet) by a plan of the other order
    from slchannel_amd.plan import HipPlan

    other = 4 if args.order == 6 else 6
    p2 = HipPlan(n, n, nz, flow.Lx, flow.Ly, flow.Lz, flow.z_f, flow.z_c, flow.gamma,
                 stretching_type=flow.stretching_type, order=other, nu=flow.nu,
                 device=flow.device)
    p2.gather("u", flow.u_nm1, xd, yd, zd, out)
    torch.cuda.synchronize()
    p2.profile(True)
    for _ in range(5):
        p2.gather("u", flow.u_nm1, xd, yd, zd, out)
    c2, t2 = p2.profile_report()["gather_u"]
    p2.profile(False)
    gather_by_order[str(other)] = round(40.0 * N_u / (t2 / c2 * 1e-3) / 1e9, 1)
    del p2

    # correctness guards on the measured run (no host sync inside the loop)
    div = flow.divergence().abs().max().item()
    ubv = ub.item()
    fp32_err = None
    if twin is not None:
        for _ in range(args.warmup + args.steps):
            twin.step_sl_bdf2(dt)
        fp32_err = {k: (a - b).abs().max().item() / b.abs().max().item()
                    for k, a, b in (("u", flow.u, twin.u), ("v", flow.v, twin.v),
                                    ("w", flow.w, twin.w), ("p", flow.p, twin.p))}
        fp32_err["steps"] = args.warmup + args.steps

    traffic = measured_traffic(lab, n, nz, args.order) if args.storage == 64 else None
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(flow, dt)

    kernels = {k: {"launches": c, "avg_ms": round(t / c, 4)} for k, (c, t) in sorted(prof.items())}
    line = {
        "metric": ("steps/s (SL-BDF2 step, 256^3 triquintic fp64)" if args.storage == 64 else
                   "steps/s (SL-BDF2 step, 256^3 triquintic, FP32-storage gathers)")
                  + ("" if args.traj_order == 2 else ", tricubic trajectories"),
        "value": round(args.steps / (t1 - t0), 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.storage == 64 else "f64 (gathered fields stored f32)",
        "data": "synthetic (seeded solenoidal k^-5/3 perturbation on Poiseuille, projected)",
        "config": {"workload": f"SL-BDF2 channel {n}x{n}x{nz} order {args.order} "
                               f"Re_b {Re_b} dt+ {dtp}",
                   "nx": n, "ny": n, "nz": nz, "interp_order": args.order, "dt": dt,
                   "storage_bits": args.storage, "traj_interp_order": args.traj_order},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1),
                     "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": traffic,
                     "bytes_per_launch": bytes_far, "avg_ms": round(avg_ms, 4),
                     "fp64_check": fp64_check(lab, avg_ms, N_u, args.order)
                     if args.traj_order == 2 else None},
        "gather_GBps": round(gather_gbps, 1),
        "gather_frac_hbm": round(gather_gbps / PEAK_HBM_GBPS, 4),
        "gather_GBps_by_order": gather_by_order,
        "max_div": div,
        "u_bulk": ubv,
        "kernels_ms": kernels,
        "fp32_storage_rel_err_vs_fp64": fp32_err,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def imex_bench(args, flow, dt_sl, cfl_imex=0.28):
    """The Eulerian IMEX step (solver.py:1055-1127) on the same state, at the
    CFL of the reference's Eulerian configuration (configs/
    re180_eulerian_reference.yaml: CFL_target 0.28), beside the SL-BDF2 step
    at its dt+ = 0.2, as simulated time per wall second."""
    import torch

    from slchannel_amd.diagnostics import cfl_dt

    n, nz = flow.nx, flow.nz
    u0, v0, w0, p0 = (t.clone() for t in (flow.u, flow.v, flow.w, flow.p))
    dt = cfl_dt(flow, cfl_imex)
    # the SL step for the comparison (same state)
    for _ in range(args.warmup):
        flow.step_sl_bdf2(dt_sl)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ks = max(5, args.steps // 2)
    for _ in range(ks):
        flow.step_sl_bdf2(dt_sl)
    torch.cuda.synchronize()
    sl_ms = (time.perf_counter() - t0) * 1e3 / ks
    flow.set_fields(u0, v0, w0, p0)
    for _ in range(args.warmup):
        flow.step_imex(dt)
    torch.cuda.synchronize()
    plan = flow.plan
    plan.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ub, _ = flow.step_imex(dt)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = plan.profile_report()
    plan.profile(False)
    ms_step = (t1 - t0) * 1e3 / args.steps
    cnt, tot = prof["imex_rhs"]
    avg_ms = tot / cnt
    # algorithmic bytes: u, v, w read once, three rhs arrays written once
    nelem = ((n + 1) * (n + 2) * (nz + 2) + (n + 2) * (n + 1) * (nz + 2) +
             (n + 2) * (n + 2) * (nz + 1))
    bytes_rhs = 2.0 * 8.0 * nelem
    achieved = bytes_rhs / (avg_ms * 1e-3) / 1e9
    div = flow.divergence().abs().max().item()
    kernels = {k: {"launches": c, "avg_ms": round(t / c, 4)} for k, (c, t) in sorted(prof.items())}
    line = {
        "metric": "steps/s (Eulerian IMEX step, 256^3 fp64)",
        "value": round(args.steps / (t1 - t0), 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded solenoidal k^-5/3 perturbation on Poiseuille, projected)",
        "config": {"workload": f"Eulerian IMEX channel {n}x{n}x{nz} CFL {cfl_imex}",
                   "nx": n, "ny": n, "nz": nz, "dt": dt},
        "roofline": {"bound": "hbm", "kernel": "imex_rhs_kernel (advection + xy-diffusion)",
                     "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": None,
                     "bytes_per_launch": bytes_rhs, "avg_ms": round(avg_ms, 4)},
        "scheme_comparison": {
            "imex_dt": dt, "imex_cfl": cfl_imex, "imex_ms_per_step": round(ms_step, 4),
            "imex_sim_time_per_s": round(dt / (ms_step * 1e-3), 5),
            "sl_dt": dt_sl, "sl_ms_per_step": round(sl_ms, 4),
            "sl_sim_time_per_s": round(dt_sl / (sl_ms * 1e-3), 5),
            "sl_over_imex": round((dt_sl / sl_ms) / (dt / ms_step), 3)},
        "max_div": div,
        "u_bulk": ub.item(),
        "kernels_ms": kernels,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def stats_bench(args):
    """On-device turbulence statistics (TurbulenceStats.accumulate_statistics,
    include/slc_stats.h) at n^3 with the primary z+ = 15 plane pair and three
    extra spectra planes (z+ = 5, 30, 100): snapshots per second, fields
    resident in HBM.  Algorithmic bytes per snapshot: the two plane passes
    read u, v, w once each (2 x 24 B per grid point)."""
    import numpy as np
    import torch

    from slchannel_amd.grid import generate_grid
    from slchannel_amd.turbstats import TurbulenceStats

    n = args.n
    nz = args.nz or n
    Lx, Ly, Lz = 4 * math.pi, 2 * math.pi, 2.0
    z_f, z_c, dz_f, dz_c = generate_grid(1.5, nz, Lz, stretching_type="symmetric")
    nu = 1.0 / 5600.0
    ts = TurbulenceStats(n, n, nz, Lx, Ly, Lz, z_c, z_f, dz_c, dz_f, Lx / n, Ly / n, nu, 180.0,
                         device="cuda", spectra_z_planes=[5.0, 30.0, 100.0])
    g = torch.Generator(device="cuda").manual_seed(1)
    f64 = dict(dtype=torch.float64, device="cuda", generator=g)
    u = 1.0 + 0.1 * torch.randn(n + 1, n + 2, nz + 2, **f64)
    v = 0.1 * torch.randn(n + 2, n + 1, nz + 2, **f64)
    w = 0.1 * torch.randn(n + 2, n + 2, nz + 1, **f64)
    for _ in range(args.warmup):
        ts.accumulate_statistics(u, v, w)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        ts.accumulate_statistics(u, v, w)
    e1.record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = e0.elapsed_time(e1) / args.steps
    nfield = ((n + 1) * (n + 2) * (nz + 2) + (n + 2) * (n + 1) * (nz + 2) +
              (n + 2) * (n + 2) * (nz + 1))
    bytes_snap = 2.0 * 8.0 * nfield
    achieved = bytes_snap / (ms * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        from oracle.turbstats import StatsOracle

        orc = StatsOracle(n, n, nz, planes=[(ts.k_bot, ts.k_top)] +
                          [(p["k_bot"], p["k_top"]) for p in ts.extra_planes])
        uh, vh, wh = (np.ascontiguousarray(t.cpu().numpy()) for t in (u, v, w))
        c0 = time.perf_counter()
        orc.accumulate(uh, vh, wh)
        c1 = time.perf_counter()
        cpu = {"value": round(1.0 / (c1 - c0), 4), "unit": "snapshots/s", "cores": 1,
               "kind": "port", "sample": f"one snapshot at {n}x{n}x{nz} (numpy oracle, "
                                         "oracle/turbstats.py)"}
    line = {
        "metric": f"snapshots/s (turbulence statistics accumulate, {n}x{n}x{nz} fp64)",
        "value": round(args.steps / (t1 - t0), 3),
        "unit": "snapshots/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random fields)",
        "config": {"workload": f"TurbulenceStats.accumulate_statistics {n}x{n}x{nz}, "
                               "4 spectra plane pairs", "nx": n, "ny": n, "nz": nz},
        "roofline": {"bound": "hbm", "kernel": "slc_stats_accumulate (two plane passes, "
                                               "reduces, spectra)",
                     "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": None,
                     "bytes_per_launch": bytes_snap, "avg_ms": round(ms, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def large_box(args):
    """1024 x 1024 x 512, Lx = 8 pi, Ly = 3 pi, tanh 1.8, Re_b = 20000, order 6,
    k^-5/3 perturbation (u'rms 0.12), dt for trajectory CFL 4; fields resident.
    Also: the isolated order-6 gather of u with feet displaced uniformly in
    +-5 cells in x and y (seeded), z at the arrival point."""
    import torch

    from slchannel_amd import SLChannelFlow
    from slchannel_amd.diagnostics import cfl_dt

    nx, ny, nz = 1024, 1024, 512
    cfg = {
        "grid": {"nx": nx, "ny": ny, "nz": nz},
        "domain": {"Lx": 8 * math.pi, "Ly": 3 * math.pi, "Lz": 2.0},
        "flow": {"Re": 20000.0, "U_bulk": 1.0, "gamma": 1.8},
        "sl": {"interp_order": args.order},
    }
    flow = SLChannelFlow(config=cfg, device="cuda:0")
    flow.initialize(seed=3000, urms=0.12, kmin=1.0, kmax=32.0, spectrum="k-5/3")
    dt = cfl_dt(flow, 4.0)
    for _ in range(args.warmup):
        flow.step_sl_bdf2(dt)
    torch.cuda.synchronize()
    plan = flow.plan
    plan.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ub, _ = flow.step_sl_bdf2(dt)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = plan.profile_report()
    plan.profile(False)
    N_u = nx * ny * nz
    lab, kname, avg_ms, bytes_k = roofline_kernel(prof, N_u, 64)
    achieved = bytes_k / (avg_ms * 1e-3) / 1e9
    # isolated gather, feet = arrival + U(-5, 5) cells in x and y
    g = torch.Generator(device="cuda").manual_seed(33)
    ii = torch.arange(N_u, device="cuda", dtype=torch.int64)
    i = (ii // (ny * nz)).double() + 1.0
    j = ((ii // nz) % ny).double() + 1.0
    xd = (i + 10.0 * torch.rand(N_u, generator=g, device="cuda", dtype=torch.float64) - 5.0) \
        * flow.dx
    yd = (j - 0.5 + 10.0 * torch.rand(N_u, generator=g, device="cuda", dtype=torch.float64)
          - 5.0) * flow.dy
    zd = flow.z_c[1:nz + 1].to("cuda").repeat(nx * ny)
    del ii, i, j
    out = torch.empty(N_u, dtype=torch.float64, device="cuda")
    plan.gather("u", flow.u, xd, yd, zd, out)
    torch.cuda.synchronize()
    plan.profile(True)
    for _ in range(3):
        plan.gather("u", flow.u, xd, yd, zd, out)
    gp = plan.profile_report()
    plan.profile(False)
    g_cnt, g_tot = gp["gather_u"]
    gather_gbps = 40.0 * N_u / (g_tot / g_cnt * 1e-3) / 1e9
    div = flow.divergence().abs().max().item()
    line = {
        "metric": "steps/s (SL-BDF2 step, 1024x1024x512 triquintic fp64)",
        "value": round(args.steps / (t1 - t0), 4), "unit": "steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round((t1 - t0) * 1e3 / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded solenoidal k^-5/3 perturbation on Poiseuille, projected)",
        "config": {"workload": "SL-BDF2 channel 1024x1024x512 order 6 Re_b 20000 CFL 4",
                   "nx": nx, "ny": ny, "nz": nz, "interp_order": args.order, "dt": dt},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1),
                     "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": None,
                     "bytes_per_launch": bytes_k, "avg_ms": round(avg_ms, 4),
                     "fp64_check": fp64_check(lab, avg_ms, N_u, args.order)},
        "gather_random_pm5_GBps": round(gather_gbps, 1),
        "max_div": div, "u_bulk": ub.item(),
        "kernels_ms": {k: {"launches": c, "avg_ms": round(t / c, 4)}
                       for k, (c, t) in sorted(prof.items())},
        "device_mem_GB": round(torch.cuda.max_memory_allocated() / 1e9, 1),
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


ENSEMBLE_GRIDS = ((64, 48, 64), (96, 64, 96), (128, 96, 128), (192, 128, 128))
SMALL_GRIDS = ((32, 24, 32), (48, 32, 48), (32, 32, 24))


def _member_cfg(nx, ny, nz, Re, order, fp32=False):
    return {"grid": {"nx": nx, "ny": ny, "nz": nz},
            "domain": {"Lx": 4 * math.pi, "Ly": 2 * math.pi, "Lz": 2.0},
            "flow": {"Re": Re, "U_bulk": 1.0, "gamma": 1.5},
            "sl": {"interp_order": order, "interp_dtype": "fp32_accum64" if fp32 else "fp64"}}


def ensemble_bench(args):
    """48 members drawn from {64x48x64, 96x64x96, 128x96x128, 192x128x128},
    Re ~ U[2800, 6000], seeds 4000 + m, order 6, dt for trajectory CFL 3 on
    each member's initial state.  Times K steps (after W warm-up steps, the
    first a BDF1 bootstrap) of (a) the plain per-run path, each member's
    SLChannelFlow stepped alone, one member after the other, and (b) the
    batched ensemble (autotuned tile shape), eager and hipGraph replay.  Also
    the FP32-storage ensemble's per-field difference to the FP64 one after
    W + K steps."""
    import numpy as np
    import torch

    from slchannel_amd import SLChannelFlow
    from slchannel_amd.diagnostics import cfl_dt
    from slchannel_amd.ensemble import OPT_GRAPH, Ensemble

    grids = ENSEMBLE_GRIDS if args.ensemble_grids == "workload" else SMALL_GRIDS
    rng = np.random.Generator(np.random.PCG64(4000))
    specs = []
    for m in range(args.members):
        nx, ny, nz = grids[int(rng.integers(len(grids)))]
        specs.append((nx, ny, nz, float(rng.uniform(2800.0, 6000.0)), 4000 + m))
    inits, flows, dts = [], [], []
    for nx, ny, nz, Re, seed in specs:
        f = SLChannelFlow(config=_member_cfg(nx, ny, nz, Re, args.order), device="cuda:0")
        f.initialize(seed=seed, urms=0.1, kmax=8.0)
        dts.append(cfl_dt(f, 3.0))
        inits.append((f.u.clone(), f.v.clone(), f.w.clone(), f.p.clone()))
        flows.append(f)
    points = sum(nx * ny * nz for nx, ny, nz, _, _ in specs)
    K, W = args.steps, args.warmup

    # (a) plain per-run path, one member at a time
    torch.cuda.synchronize()
    plain_s = 0.0
    for f, dt in zip(flows, dts):
        for _ in range(W):
            f.step_sl_bdf2(dt)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(K):
            f.step_sl_bdf2(dt)
        torch.cuda.synchronize()
        plain_s += time.perf_counter() - t
    plain = args.members * K / plain_s

    # (b) batched ensemble
    def build(storage):
        ens = Ensemble(order=args.order, storage=storage)
        if args.bdf2_split:
            from slchannel_amd.ensemble import OPT_BDF2_FUSED
            ens.set_option(OPT_BDF2_FUSED, 0)
        ids = [ens.add(_member_cfg(nx, ny, nz, Re, args.order, storage == 32), dt, fields=ini)
               for (nx, ny, nz, Re, _), dt, ini in zip(specs, dts, inits)]
        return ens, ids

    ens, ids = build(64)
    for _ in range(W):
        ens.step()
    tune, tune_grid = ens.autotune(reps=2)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(K):
        ens.step()
    torch.cuda.synchronize()
    eager_s = time.perf_counter() - t
    ens.set_option(OPT_GRAPH, 1)
    ens.step()  # capture
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(K):
        ens.step()
    torch.cuda.synchronize()
    graph_s = time.perf_counter() - t
    batched = args.members * max(K / eager_s, K / graph_s)
    # bit-identity spot check against the plain path after W + K steps (member 0)
    f0 = ens.get_fields(ids[0])  # member 0 is now W + 2K + 1 steps in; step the plain one on
    for _ in range(K + 1):
        flows[0].step_sl_bdf2(dts[0])
    same0 = all(torch.equal(f0[k], getattr(flows[0], k)) for k in "uvwp")

    # FP32-storage ensemble vs FP64 ensemble after W + K steps
    e64, i64 = build(64)
    e32, i32 = build(32)
    for _ in range(W + K):
        e64.step()
        e32.step()
    err = {k: 0.0 for k in "uvwp"}
    for a, b in zip(i64, i32):
        fa, fb = e64.get_fields(a), e32.get_fields(b)
        for k in "uvwp":
            err[k] = max(err[k], (fb[k] - fa[k]).abs().max().item() / fa[k].abs().max().item())
    line = {
        "metric": f"member-steps/s ({args.members}-member uneven SL-BDF2 ensemble, triquintic fp64)",
        "value": round(batched, 2), "unit": "member-steps/s", "n_gpus": 1, "steps": K,
        "warmup": W, "ms_per_step": round(min(eager_s, graph_s) * 1e3 / K, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded solenoidal perturbations on Poiseuille, projected)",
        "config": {"workload": f"{args.members} members from {list(grids)}, "
                               "Re U[2800,6000], order 6, CFL 3",
                   "members": args.members, "grid_points": points},
        "plain_per_run_member_steps_per_s": round(plain, 2),
        "batched_vs_plain": round(batched / plain, 3),
        "eager_member_steps_per_s": round(args.members * K / eager_s, 2),
        "graph_member_steps_per_s": round(args.members * K / graph_s, 2),
        "grid_point_steps_per_s": round(points * K / min(eager_s, graph_s), 1),
        "autotune_ms": {"tile_shape": tune, "grid_per_cu": tune_grid},
        "member0_identical_to_plain": bool(same0),
        "fp32_storage_rel_err_vs_fp64": {k: float(f"{v:.3e}") for k, v in err.items()},
        "fp32_storage_err_after_steps": W + K,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def serve_bench(args):
    """Poisson arrivals (rate per second, seeded) of requests with grids from
    {64^3, 128x96x128, 256^3, 384x384x256}, Re ~ U[2800, 10000], N ~ U[20, 200]
    steps, order 6, dt for trajectory CFL 3; admission by HBM footprint under
    the given budget.  The plain per-run path is each request run alone
    (SLChannelFlow: its seeded initial state, then its steps; solver plans
    built beforehand, as the server's per-grid initialisers are), served
    first-come-first-served under the same arrival times (its queue simulated
    from the measured stand-alone durations)."""
    import numpy as np

    from slchannel_amd.serve import Server, plain_fifo, poisson_requests, run_alone

    grids = ((64, 64, 64), (128, 96, 128), (256, 256, 256), (384, 384, 256))
    if args.serve_grids == "small":
        grids = SMALL_GRIDS
    reqs = poisson_requests(args.requests, args.rate, grids, (2800.0, 10000.0), (20, 200),
                            seed=5000, order=args.order)
    srv = Server(hbm_budget=int(args.hbm_budget_gb * 1e9))
    srv.warmup([r.config for r in reqs])  # server start-up: one initialiser per grid
    results, m = srv.run(reqs)
    durations, plain_lat = [], []
    for r in reqs:
        res, el = run_alone(r, keep_fields=False)
        durations.append(el)
        plain_lat += res.latencies_ms
    pf = plain_fifo(reqs, durations)
    total_steps = sum(r.n_steps for r in reqs)
    plain = pf["member_steps_per_s"]
    line = {
        "metric": "member-steps/s served (continuous batching, Poisson arrivals)",
        "value": round(m["member_steps_per_s"], 2), "unit": "member-steps/s", "n_gpus": 1,
        "steps": total_steps, "warmup": 0, "ms_per_step": round(m["p50_step_ms"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded solenoidal perturbations on Poiseuille, projected)",
        "config": {"workload": f"{args.requests} requests, rate {args.rate}/s, grids "
                               f"{list(grids)}, Re U[2800,10000], N U[20,200]",
                   "hbm_budget_GB": args.hbm_budget_gb},
        "p50_step_ms": round(m["p50_step_ms"], 3), "p99_step_ms": round(m["p99_step_ms"], 3),
        "p50_time_in_system_s": round(m["p50_time_in_system_s"], 3),
        "p99_time_in_system_s": round(m["p99_time_in_system_s"], 3),
        "wall_s": round(m["wall_s"], 2),
        "device_step_s": round(m["device_step_s"], 2),
        "host_s": m["host_s"],
        "plain_fifo_member_steps_per_s": round(plain, 2),
        "plain_p50_step_ms": round(float(np.percentile(plain_lat, 50)), 3),
        "plain_p99_step_ms": round(float(np.percentile(plain_lat, 99)), 3),
        "plain_p50_time_in_system_s": round(pf["p50_time_in_system_s"], 3),
        "plain_p99_time_in_system_s": round(pf["p99_time_in_system_s"], 3),
        "plain_back_to_back_member_steps_per_s": round(total_steps / sum(durations), 2),
        "served_vs_plain": round(m["member_steps_per_s"] / plain, 3),
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def advect_bench(args):
    """The reference's published advect microbenchmark
    (benchmarks/bench_interp.py:34-111): the three-component
    SLAdvector.advect(u, v, w, u, v, w, dt) -- departure feet plus the
    order-P gather of every component -- at 768 x 768 x 180 (Lx 15.707,
    Ly 6.282, Lz 2, tanh 2.6), u = 1 + 0.1 U(0,1), v = 0.2 U(0,1),
    w = 0.1 U(0,1), dt 0.05; 3 warm-up calls, then the mean of 5.  The
    reference publishes 126 ms (order 4) and 143 ms (order 6) for its Triton
    fp32 pipeline on a GB10 (report/sl_dns_report.tex:765); here FP64 and the
    FP32-storage variant (FP64 weights and accumulation)."""
    import torch

    from slchannel_amd import SLAdvector
    from slchannel_amd.grid import generate_grid

    nx, ny, nz = 768, 768, 180
    Lx, Ly, Lz, gamma = 15.707, 6.282, 2.0, 2.6
    z_f, z_c, _, _ = generate_grid(gamma, nz, Lz)
    g = torch.Generator(device="cuda").manual_seed(7)

    def rand(*shape):
        return torch.rand(*shape, generator=g, device="cuda", dtype=torch.float64)

    u = 1.0 + 0.1 * rand(nx + 1, ny + 2, nz + 2)
    v = 0.2 * rand(nx + 2, ny + 1, nz + 2)
    w = 0.1 * rand(nx + 2, ny + 2, nz + 1)
    dt = torch.tensor(0.05, dtype=torch.float64)
    published = {4: 126.0, 6: 143.0}
    rows = []
    for order in (4, 6):
        for dtype in ("fp64", "fp32_accum64"):
            adv = SLAdvector(nx, ny, nz, Lx / nx, Ly / ny, Lx, Ly, Lz, z_f, z_c, gamma,
                             order=order, interp_dtype=dtype, device="cuda:0")
            for _ in range(3):
                adv.advect(u, v, w, u, v, w, dt)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(5):
                adv.advect(u, v, w, u, v, w, dt)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) / 5 * 1e3
            rows.append({"order": order, "interp_dtype": dtype, "ms": round(ms, 3),
                         "published_gb10_triton_fp32_ms": published[order]})
            del adv
            torch.cuda.empty_cache()
    best6 = min(r["ms"] for r in rows if r["order"] == 6 and r["interp_dtype"] == "fp64")
    line = {
        "metric": "ms per three-component advect (768x768x180, bench_interp.py workload)",
        "value": best6, "unit": "ms", "n_gpus": 1, "steps": 5, "warmup": 3,
        "ms_per_step": best6, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (the reference benchmark's uniform random fields)",
        "config": {"workload": "SLAdvector.advect 768x768x180, Lx 15.707, Ly 6.282, "
                               "tanh 2.6, dt 0.05", "nx": nx, "ny": ny, "nz": nz},
        "cases": rows,
        "note": "the published figures are the reference's Triton fp32 pipeline on a GB10; "
                "context, not a like-for-like target",
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def roofline_kernel(prof, n_pts, storage):
    """The dominant kernel of the SL-BDF2 step: the fused BDF2 pass of u
    (default), else the far pass of u.  Returns (label, kernel name, average
    launch ms, algorithmic bytes per launch).  Algorithmic bytes per arrival
    point, each array counted once: fused -- U* 3x8 + V^{n-1} 8 + V^n 8 read
    (each gathered and differenced), the prediction 8 written (float storage:
    + 4 + 4 for the gathered float copies); far -- U* 3x8 + V^{n-1} 8 read,
    the far value 8 + the near foot 3x8 written (float storage: the gathered
    field 4 instead of 8)."""
    fused = "bdf2_u" in prof
    lab = "bdf2_u" if fused else "bdf2_far_u"
    cnt, tot = prof[lab]
    if fused:
        per_pt = 48.0 if storage == 64 else 56.0
    else:
        per_pt = 64.0 if storage == 64 else 60.0
    kname = "sl_tile_kernel<u,P,%s> (%s)" % ("BDF2" if fused else "FAR", lab)
    return lab, kname, tot / cnt, per_pt * n_pts


PEAK_FP64_TFLOPS = 78.6


def fp64_flops_per_point(order, fused):
    """FP64 flops per arrival point of the dominant SL kernel, counted from
    the algorithm (FMA = 2).  Gather at order P: the staged contraction
    2 (P^3 + P^2 + P), Lagrange weights P^2 per uniform dimension (P
    differences, P products of P - 1 factors and a scale) and P^2 for z.
    Tracer: 9 trilinear samples of U* (arrival point + one midpoint per foot
    at n_traj_iters = 2; 41 flops each: coordinates, 2-point weights, eight
    weighted terms) and 52 flops of midpoints, feet and stencil origins.
    Fused BDF2 pass: two gathers, the tracer, the two xy-RHS stencils and the
    BDF2 combination (27).  Far pass: one gather and the tracer.  (At order 6
    the fused count is 1696; the SQ_INSTS_VALU_FLOPS_FP64 counter of the same
    launch gives 1750.)"""
    P = order
    gather = 2 * (P**3 + P**2 + P) + 3 * P**2
    tracer = 9 * 41 + 52
    return 2 * gather + tracer + 27 if fused else gather + tracer


def fp64_check(prof_label, avg_ms, n_pts, order):
    """The HBM roofline of the SL pass checked against the FP64 VALU peak."""
    flops = fp64_flops_per_point(order, prof_label == "bdf2_u") * n_pts
    tf = flops / (avg_ms * 1e-3) / 1e12
    return {"flops_per_launch": flops, "achieved": round(tf, 2), "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s", "frac": round(tf / PEAK_FP64_TFLOPS, 4)}


def measured_traffic(label, n, nz, order):
    """HBM bytes per launch of the roofline kernel (launch label ``label``)
    from the committed PMC summary (profiles/traffic_<label>_*.json,
    FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel calibration), if one
    exists for this workload; else None."""
    path = os.path.join(REPO, "profiles", f"traffic_{label}_{n}x{n}x{nz}_o{order}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def cpu_baseline(flow, dt):
    """One BDF2 step of the CPU oracle on the GPU's current 256^3 state."""
    import numpy as np

    from oracle import oracle as orc

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    co = orc.ChannelOracle(flow.nx, flow.ny, flow.nz, flow.Lx, flow.Ly, flow.Lz, flow.gamma,
                           1.0 / flow.nu, order=flow.sl_order, z_f=flow.z_f.numpy(),
                           z_c=flow.z_c.numpy(), tanh_g=flow.plan.tanh_g)
    co.set_fields(*(t.cpu().numpy() for t in (flow.u, flow.v, flow.w, flow.p)))
    co.set_history(*(t.cpu().numpy() for t in (flow.u_nm1, flow.v_nm1, flow.w_nm1)), dt)
    t0 = time.perf_counter()
    co.step_sl_bdf2(dt)
    sec = time.perf_counter() - t0
    del np
    return {"value": round(1.0 / sec, 5), "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"one BDF2 step of the CPU oracle at {flow.nx}x{flow.ny}x{flow.nz} "
                      f"(numpy + OpenMP C, {threads} threads; {sec:.1f} s)"}


if __name__ == "__main__":
    main()
