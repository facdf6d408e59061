"""Serving loop: continuous batching of independent channel runs on one GPU.

Requests (a channel config, a seed, a number of steps, an arrival time)
arrive over time.  The loop admits arrived requests in FIFO order while the
ensemble's HBM footprint plus the newcomer's (slc_ens_member_bytes) stays
within the budget, initialises them on the device, and advances every
admitted run by one SL-BDF2 step per iteration with the batched ensemble
(one launch per substep for all runs, include/slc_ens.h).  Runs join and
leave between steps: a run admitted while others are mid-flight bootstraps
with BDF1 inside the same launches in which the others take BDF2 steps, and
a run that has done its steps is copied out and its memory released, so a
waiting request can take its place.

Reported per run: admission and completion times, every step's latency (the
wall time of the batched step it took part in, launch to completion), and
optionally its final fields.  Aggregates: p50 / p99 step latency and member
steps per second over the busy period.

The reference has no serving layer (one run per process, solver.py:1280-1500).
"""

import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .diagnostics import cfl_dt
from .ensemble import Ensemble, _INIT_CACHE, _config_grid, initial_fields, member_bytes


@dataclass
class Request:
    rid: int
    config: dict
    n_steps: int
    seed: int
    arrival: float = 0.0      # seconds after the loop starts
    dt: float = None          # None: trajectory CFL `cfl` on the initial state
    cfl: float = 3.0
    urms: float = 0.1
    kmax: float = 8.0


@dataclass
class Result:
    rid: int
    admitted: float = 0.0
    finished: float = 0.0
    dt: float = 0.0
    steps: int = 0
    latencies_ms: list = field(default_factory=list)
    fields: dict = None
    stats: list = None


def _prepare(req, device):
    """Initial state and dt of a request on the device (SLChannelFlow.initialize)."""
    u, v, w, p = initial_fields(req.config, req.seed, urms=req.urms, kmax=req.kmax,
                                device=device)
    dt = req.dt
    if dt is None:
        c = _config_grid(req.config)
        key = (c["nx"], c["ny"], c["nz"], c["Lx"], c["Ly"], c["Lz"], c["gamma"], c["stretching"],
               c["top"], c["order"], str(device))
        dt = cfl_dt(_INIT_CACHE[key], req.cfl)
    return (u, v, w, p), float(dt)


class Server:
    """Continuous-batching loop over one Ensemble."""

    def __init__(self, hbm_budget=None, max_members=512, order=6, storage=64, graph=False,
                 keep_fields=False, device=None):
        self.ens = Ensemble(order=order, storage=storage, device=device, graph=graph)
        self.device = self.ens.device
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(0.6 * free)
        self.budget = int(hbm_budget)
        self.max_members = int(max_members)
        self.keep_fields = keep_fields
        self.storage = storage

    def fits(self, req):
        need = member_bytes(req.config, self.storage)
        return need <= self.budget - self.ens.bytes() and len(self.ens) < self.max_members

    def run(self, requests, log=None):
        """Serve `requests` (arrival times relative to the start); returns
        (results by rid, metrics dict)."""
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        for r in pending:
            if member_bytes(r.config, self.storage) > self.budget:
                raise ValueError(f"request {r.rid} alone exceeds the HBM budget")
        results = {}
        active = {}  # member id -> (request, result)
        t0 = time.perf_counter()
        busy = 0.0
        n_member_steps = 0
        first_admit = None
        while pending or active:
            now = time.perf_counter() - t0
            # admission: FIFO over the requests that have arrived
            while pending and pending[0].arrival <= now and self.fits(pending[0]):
                req = pending.pop(0)
                fields, dt = _prepare(req, self.device)
                mid = self.ens.add(req.config, dt, fields=fields)
                res = Result(rid=req.rid, admitted=time.perf_counter() - t0, dt=dt)
                active[mid] = (req, res)
                first_admit = res.admitted if first_admit is None else first_admit
                if log:
                    log(f"admit {req.rid} {_config_grid(req.config)['nx']}x"
                        f"{_config_grid(req.config)['ny']}x{_config_grid(req.config)['nz']} "
                        f"steps={req.n_steps} active={len(active)}")
            if not active:
                if pending:
                    wait = pending[0].arrival - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.05))
                continue
            torch.cuda.synchronize(self.device)
            ts = time.perf_counter()
            self.ens.step()
            torch.cuda.synchronize(self.device)
            te = time.perf_counter()
            lat = (te - ts) * 1e3
            busy += te - ts
            n_member_steps += len(active)
            done = []
            for mid, (req, res) in active.items():
                res.latencies_ms.append(lat)
                res.steps += 1
                if res.steps >= req.n_steps:
                    done.append(mid)
            for mid in done:
                req, res = active.pop(mid)
                if self.keep_fields:
                    res.fields = self.ens.get_fields(mid)
                res.stats = self.ens.stats(mid).cpu().tolist()
                res.finished = time.perf_counter() - t0
                self.ens.remove(mid)
                results[req.rid] = res
        t_end = time.perf_counter() - t0
        lats = np.array([x for r in results.values() for x in r.latencies_ms])
        span = t_end - (first_admit or 0.0)
        metrics = {
            "requests": len(results),
            "member_steps": int(n_member_steps),
            "p50_step_ms": float(np.percentile(lats, 50)) if lats.size else math.nan,
            "p99_step_ms": float(np.percentile(lats, 99)) if lats.size else math.nan,
            "member_steps_per_s": n_member_steps / span if span > 0 else math.nan,
            "member_steps_per_busy_s": n_member_steps / busy if busy > 0 else math.nan,
            "wall_s": t_end,
            "busy_s": busy,
        }
        return results, metrics


def poisson_requests(n, rate, grids, re_range, steps_range, seed, order=6, base_seed=5000):
    """n requests with exponential inter-arrival times (rate per second),
    grid sizes, Re and step counts drawn uniformly (seeded)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        nx, ny, nz = grids[int(rng.integers(len(grids)))]
        Re = float(rng.uniform(*re_range))
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
        cfg = {"grid": {"nx": nx, "ny": ny, "nz": nz},
               "domain": {"Lx": 4 * math.pi, "Ly": 2 * math.pi, "Lz": 2.0},
               "flow": {"Re": Re, "U_bulk": 1.0, "gamma": 1.5},
               "sl": {"interp_order": order}}
        out.append(Request(rid=i, config=cfg, n_steps=steps, seed=base_seed + i, arrival=t))
    return out


def run_alone(req, device=None, keep_fields=True):
    """The plain per-run path: one SLChannelFlow stepping this request alone
    (one synchronisation at the end).  Returns (Result, seconds of stepping)."""
    from .solver import SLChannelFlow

    flow = SLChannelFlow(config=req.config, device=device)
    fields, dt = _prepare(req, flow.device)
    flow.set_fields(*fields)
    torch.cuda.synchronize(flow.device)
    res = Result(rid=req.rid, dt=dt)
    t = time.perf_counter()
    for _ in range(req.n_steps):
        ub, fo = flow.step_sl_bdf2(dt)
    torch.cuda.synchronize(flow.device)
    el = time.perf_counter() - t
    res.steps = req.n_steps
    res.latencies_ms = [el * 1e3 / req.n_steps] * req.n_steps
    if keep_fields:
        res.fields = {"u": flow.u, "v": flow.v, "w": flow.w, "p": flow.p}
    res.stats = [ub.item(), fo.item(), float(flow.sl.n_clamped_last.item()), float(res.steps)]
    return res, el
