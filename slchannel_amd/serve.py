"""Serving loop: continuous batching of independent channel runs on one GPU.

Requests (a channel config, a seed, a number of steps, an arrival time)
arrive over time.  The loop initialises admitted requests on the device and
advances every admitted run by one SL-BDF2 step per iteration with the
batched ensemble (one launch per substep for all runs, include/slc_ens.h).
Admission needs the ensemble's HBM footprint plus the newcomer's
(slc_ens_member_bytes) to stay within the budget, and follows the device's
compute: the oldest waiting request joins when the batch is empty or its
grid points fit the batch budget (`batch_points`), so runs that each fill
the GPU take it one after the other at full speed, in arrival order; and any
arrived small request (at most `small_points` grid points) joins at once
while the small runs' points fit the budget -- it rides in the launches of
whatever runs, costing its own arithmetic and no launches of its own.  Runs join and
leave between steps: a run admitted while others are mid-flight bootstraps
with BDF1 inside the same launches in which the others take BDF2 steps, and
a run that has done its steps is copied out and its memory released, so a
waiting request can take its place.

Reported per run: arrival, admission and completion times, every step's
latency (the device time of the batched step it took part in, from
hipEvents), its final run diagnostics (max|div|, u_tau, CFL, non-finite
entries; slc_ens_diagnostics), a status, and optionally its final fields.
Aggregates: p50 / p99 step latency, p50 / p99 time in system, member steps
per second over the span from the first admission to the last completion.

Guards (the reference run loop's, solver.py:1394-1479, per run): a request
whose given initial state holds NaN/Inf is rejected without being stepped;
with ``guard_every = k`` every run's diagnostics are taken every k of its
steps into pinned host memory and read once their event has completed (no
host synchronisation), and a run that turned non-finite, or whose u_tau
exceeded the blow-up factor on three consecutive checks, is stopped and its
slot released.

The reference has no serving layer (one run per process, solver.py:1280-1500).
"""

import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .diagnostics import BlowupGuard
from .ensemble import N_STATS, Ensemble, _INIT_CACHE, _config_grid, initial_fields, member_bytes


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
    fields: tuple = None      # (u, v, w[, p]) device tensors: initial state instead of `seed`
    u_tau_nominal: float = None  # blow-up guard reference (None: non-finite check only)


@dataclass
class Result:
    rid: int
    arrival: float = 0.0
    admitted: float = 0.0
    finished: float = 0.0
    dt: float = 0.0
    steps: int = 0
    latencies_ms: list = field(default_factory=list)
    fields: dict = None
    stats: list = None
    diag: list = None         # [max|div|, u_tau, CFL, dti, non-finite] at completion
    # ok | rejected_invalid (diag = [reason]) | rejected_nonfinite |
    # stopped_nonfinite | stopped_blowup
    status: str = "pending"


def _grid_key(config, device=None):
    c = _config_grid(config)
    key = (c["nx"], c["ny"], c["nz"], c["Lx"], c["Ly"], c["Lz"], c["gamma"], c["stretching"],
           c["top"], c["order"])
    return key if device is None else key + (str(device),)


def _prepare_async(req, device):
    """Initial state of a request on the device (SLChannelFlow.initialize, or
    the request's own fields), launched on the current stream without a host
    synchronisation.  Returns (fields, dt, probe): dt is the request's own dt
    or None; probe is a 2-double device tensor [dti, number of non-finite
    entries] (dti: the trajectory CFL inverse time-scale, operators.py:839-940,
    via slc_diagnostics) for the caller to read once the stream is done."""
    probe = torch.zeros(2, dtype=torch.float64, device=device)
    if req.fields is not None:
        if req.dt is None:
            raise ValueError(f"request {req.rid}: explicit fields need an explicit dt")
        u, v, w = (t.to(device=device, dtype=torch.float64).contiguous() for t in req.fields[:3])
        p = req.fields[3] if len(req.fields) > 3 and req.fields[3] is not None else None
        if p is not None:
            p = p.to(device=device, dtype=torch.float64).contiguous()
        bad = sum((~torch.isfinite(t)).sum() for t in (u, v, w, p) if t is not None)
        probe[1:2].copy_(bad.to(torch.float64).reshape(1))
        return (u, v, w, p), float(req.dt), probe
    u, v, w, p = initial_fields(req.config, req.seed, urms=req.urms, kmax=req.kmax,
                                device=device)
    if req.dt is not None:
        return (u, v, w, p), float(req.dt), probe
    out = torch.empty(5, dtype=torch.float64, device=device)
    _INIT_CACHE[_grid_key(req.config, device)].plan.diagnostics(u, v, w, None, 0.0, out)
    probe[0:1].copy_(out[3:4])
    return (u, v, w, p), None, probe


def validate(req):
    """Why a request cannot be served (None if it can): a config the ensemble
    does not run, a non-positive step count or dt, or given fields without a
    dt or of the wrong shapes.  Checked before the loop starts, so one bad
    request is rejected on its own instead of aborting the run."""
    try:
        c = _config_grid(req.config)
    except (KeyError, TypeError, ValueError) as e:
        return f"config: {e}"
    if int(req.n_steps) < 1:
        return f"n_steps must be >= 1, got {req.n_steps}"
    if req.dt is not None and not (float(req.dt) > 0.0):
        return f"dt must be positive, got {req.dt}"
    if req.fields is not None:
        if req.dt is None:
            return "explicit fields need an explicit dt"
        from .plan import field_shapes

        shapes = field_shapes(c["nx"], c["ny"], c["nz"])
        if len(req.fields) not in (3, 4):
            return "fields must be (u, v, w) or (u, v, w, p)"
        for name, t in zip("uvwp", req.fields):
            if t is None and name == "p":
                continue
            if not isinstance(t, torch.Tensor) or not t.is_floating_point():
                return f"field {name}: a floating-point tensor is required"
            if tuple(t.shape) != shapes[name]:
                return f"field {name}: shape {tuple(t.shape)}, expected {shapes[name]}"
    return None


def init_cache_bytes(device):
    """Device bytes held by the per-grid initialisers of `device` (their plans
    and fields), which share the HBM budget with the members."""
    def tensors(obj):
        out = []
        for t in vars(obj).values():
            if isinstance(t, torch.Tensor) and t.is_cuda:
                out.append(t)
            elif isinstance(t, (list, tuple)):
                out.extend(x for x in t if isinstance(x, torch.Tensor) and x.is_cuda)
        return out

    total = 0
    for key, flow in _INIT_CACHE.items():
        if key[-1] != str(device):
            continue
        total += flow.plan.workspace_bytes()
        seen = {}
        for obj in (flow, flow.sl, flow.sl._hip):
            for t in tensors(obj):
                seen[t.data_ptr()] = t.numel() * t.element_size()
        total += sum(seen.values())
    return total


def _points(config):
    c = _config_grid(config)
    return c["nx"] * c["ny"] * c["nz"]


def _dt_of(req, dti):
    """cfl_dt's rule: dt = CFL / dti (dti > 1e-10, else CFL)."""
    return req.cfl / (dti if dti > 1e-10 else 1.0)


def _prepare(req, device):
    """_prepare_async plus the one host read of dt."""
    fields, dt, probe = _prepare_async(req, device)
    if dt is None:
        dt = _dt_of(req, float(probe[0].item()))
    return fields, float(dt)


class Server:
    """Continuous-batching loop over one Ensemble."""

    def __init__(self, hbm_budget=None, max_members=512, order=6, storage=64, graph=False,
                 keep_fields=False, device=None, guard_every=0, blowup_factor=2.0,
                 batch_points=1 << 22, small_points=None):
        self.ens = Ensemble(order=order, storage=storage, device=device, graph=graph)
        self.device = self.ens.device
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(0.6 * free)
        self.budget = int(hbm_budget)
        self.max_members = int(max_members)
        self.keep_fields = keep_fields
        self.storage = storage
        self.guard_every = int(guard_every)
        self.blowup_factor = float(blowup_factor)
        # compute-aware admission: grid points of the batch beyond which one
        # step already fills the device, and the size of a rider
        self.batch_points = int(batch_points)
        self.small_points = int(small_points if small_points is not None else batch_points // 2)
        # admissions (initial fields, dt) run on one stream per grid size, so
        # they overlap the batched steps instead of queueing behind them, and
        # two admissions of one grid (which share its initialiser's plan)
        # stay ordered
        self._streams = {}

    def _stream_for(self, config):
        key = _grid_key(config)
        st = self._streams.get(key)
        if st is None:
            st = self._streams[key] = torch.cuda.Stream(self.device)
        return st

    def warmup(self, configs):
        """Server start-up for a grid catalogue: builds the per-grid initialiser
        (its rocFFT plans) of every distinct grid before requests arrive."""
        seen = set()
        for cfg in configs:
            c = _config_grid(cfg)
            key = (c["nx"], c["ny"], c["nz"], c["Lx"], c["Ly"], c["Lz"], c["gamma"],
                   c["stretching"], c["top"], c["order"])
            if key in seen:
                continue
            seen.add(key)
            with torch.cuda.stream(self._stream_for(cfg)):
                initial_fields(cfg, 0, device=self.device)
        torch.cuda.synchronize(self.device)

    def footprint(self):
        """HBM the server holds: the ensemble (live members, pooled arenas,
        class buffers) plus the per-grid initialisers."""
        return self.ens.bytes() + init_cache_bytes(self.device)

    def fits(self, req, reserved=0, preparing=0, trim=False):
        """Whether the request fits the budget beside what the server holds.
        Released arenas kept for reuse count as held; with trim=True, the
        oldest of them are freed when that makes the request fit."""
        if len(self.ens) + preparing >= self.max_members:
            return False
        need = member_bytes(req.config, self.storage)
        room = self.budget - self.footprint() - reserved
        if need <= room:
            return True
        pooled = self.ens.trim_pool(-1)
        if need > room + pooled:
            return False
        if trim:
            self.ens.trim_pool(pooled - (need - room))
        return True

    def _pick(self, pending, now, active_pts, small_pts, reserved, preparing):
        """The next request to admit (its index in pending) or None: the
        oldest arrived one when the batch is empty or it fits the batch
        budget, else the oldest arrived small one that fits the small runs'
        budget.  Both also need the HBM budget."""
        if not pending or pending[0].arrival > now:
            return None
        head = _points(pending[0].config)
        if (active_pts == 0 or active_pts + head <= self.batch_points) and \
                self.fits(pending[0], reserved, preparing, trim=True):
            return 0
        for i, r in enumerate(pending[1:], 1):
            if r.arrival > now:
                break
            p = _points(r.config)
            if p <= self.small_points and small_pts + p <= self.batch_points and \
                    self.fits(r, reserved, preparing, trim=True):
                return i
        return None

    def run(self, requests, log=None, inflight=3):
        """Serve `requests` (arrival times relative to the start); returns
        (results by rid, metrics dict).  Steps are pipelined: the host keeps
        at most `inflight` batched steps queued ahead of the device and only
        synchronises to retire a finished run; each step's latency is its
        device time from hipEvents."""
        results = {}
        pending = []
        for r in sorted(requests, key=lambda r: (r.arrival, r.rid)):
            why = validate(r)
            if why is None and member_bytes(r.config, self.storage) > self.budget:
                why = "alone exceeds the HBM budget"
            if why is None:
                pending.append(r)
            else:
                results[r.rid] = Result(rid=r.rid, arrival=r.arrival, admitted=r.arrival,
                                        finished=r.arrival, status="rejected_invalid")
                results[r.rid].diag = [why]
        active = {}  # member id -> (request, result)
        queued = []  # (start event, end event, results stepped)
        retiring = []  # (event after the copies out, request, result, pinned stats, diag)
        preparing = []  # (event, request, fields, dt, pinned probe, stream, bytes)
        guards = {}  # member id -> BlowupGuard
        checks = []  # (event, member id, rid, pinned diagnostics)
        t0 = time.perf_counter()
        n_member_steps = 0
        first_admit = None
        host = {"admit_s": 0.0, "step_s": 0.0, "retire_s": 0.0, "wait_s": 0.0}

        step_ms = []
        pts = {}  # member id -> grid points

        def drain(keep):
            while len(queued) > keep:
                ev0, ev1, stepped = queued.pop(0)
                ev1.synchronize()
                ms = ev0.elapsed_time(ev1)
                step_ms.append(ms)
                for res in stepped:
                    res.latencies_ms.append(ms)

        while pending or active or preparing:
            now = time.perf_counter() - t0
            ta = time.perf_counter()
            # admission, FIFO over the requests that have arrived: the initial
            # state is launched on the grid's stream and read back (dt, the
            # finiteness check) through pinned memory once that stream is done
            reserved = sum(x[6] for x in preparing)
            active_pts = sum(pts.values()) + sum(_points(x[1].config) for x in preparing)
            small_pts = sum(p for p in pts.values() if p <= self.small_points) + sum(
                _points(x[1].config) for x in preparing
                if _points(x[1].config) <= self.small_points)
            pick = self._pick(pending, now, active_pts, small_pts, reserved, len(preparing))
            if pick is not None:
                req = pending.pop(pick)
                st = self._stream_for(req.config)
                with torch.cuda.stream(st):
                    fields, dt, probe = _prepare_async(req, self.device)
                    h = torch.empty(2, dtype=torch.float64, pin_memory=True)
                    h.copy_(probe, non_blocking=True)
                    ev = torch.cuda.Event()
                    ev.record()
                preparing.append((ev, req, fields, dt, h, st,
                                  member_bytes(req.config, self.storage)))
            while preparing and preparing[0][0].query():
                _, req, fields, dt, h, st, _ = preparing.pop(0)
                probe = h.tolist()
                if dt is None:
                    dt = _dt_of(req, probe[0])
                if probe[1] > 0:
                    res = Result(rid=req.rid, arrival=req.arrival, dt=dt,
                                 status="rejected_nonfinite")
                    res.admitted = res.finished = time.perf_counter() - t0
                    results[req.rid] = res
                    continue
                with torch.cuda.stream(st):
                    mid = self.ens.add(req.config, dt, fields=fields)
                res = Result(rid=req.rid, admitted=time.perf_counter() - t0, dt=dt)
                res.arrival = req.arrival
                guards[mid] = BlowupGuard(req.u_tau_nominal, self.blowup_factor)
                active[mid] = (req, res)
                pts[mid] = _points(req.config)
                first_admit = res.admitted if first_admit is None else first_admit
                if log:
                    c = _config_grid(req.config)
                    log(f"admit {req.rid} {c['nx']}x{c['ny']}x{c['nz']} "
                        f"steps={req.n_steps} footprint={self.footprint()} "
                        f"active={len(active)}")
            host["admit_s"] += time.perf_counter() - ta
            if not active and not preparing and pending and pending[0].arrival <= now and \
                    not self.fits(pending[0]):
                # nothing runs that could free memory: the server's fixed
                # footprint (pooled arenas, initialisers) leaves no room
                req = pending.pop(0)
                res = Result(rid=req.rid, arrival=req.arrival, status="rejected_invalid")
                res.admitted = res.finished = time.perf_counter() - t0
                res.diag = ["does not fit the HBM budget beside the server's fixed footprint"]
                results[req.rid] = res
                continue
            if not active:
                if preparing:
                    time.sleep(1e-4)
                elif pending:
                    wait = pending[0].arrival - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.05))
                continue
            tw = time.perf_counter()
            drain(inflight - 1)
            ts = time.perf_counter()
            host["wait_s"] += ts - tw
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            ev0.record()
            self.ens.step()
            ev1.record()
            host["step_s"] += time.perf_counter() - ts
            queued.append((ev0, ev1, [res for _, res in active.values()]))
            n_member_steps += len(active)
            done = []
            for mid, (req, res) in active.items():
                res.steps += 1
                if res.steps >= req.n_steps:
                    done.append((mid, "ok"))
                elif self.guard_every and res.steps % self.guard_every == 0:
                    h = torch.empty(5, dtype=torch.float64, pin_memory=True)
                    h.copy_(self.ens.diagnostics(mid), non_blocking=True)
                    ev = torch.cuda.Event()
                    ev.record()
                    checks.append((ev, mid, req.rid, h))
            # guard results whose copies have landed (no waiting)
            while checks and checks[0][0].query():
                _, mid, rid, h = checks.pop(0)
                if mid in active and active[mid][0].rid == rid and \
                        all(m != mid for m, _ in done):
                    verdict = guards[mid].check(h.tolist())
                    if verdict is not None:
                        done.append((mid, "stopped_" + verdict))
            tr = time.perf_counter()
            for mid, status in done:
                req, res = active.pop(mid)
                guards.pop(mid, None)
                pts.pop(mid, None)
                res.status = status
                # copies are stream-ordered behind the member's last step;
                # the arena is reused only by work queued after them
                if self.keep_fields:
                    res.fields = self.ens.get_fields(mid)
                hs = torch.empty(N_STATS, dtype=torch.float64, pin_memory=True)
                hd = torch.empty(5, dtype=torch.float64, pin_memory=True)
                hs.copy_(self.ens.stats(mid), non_blocking=True)
                hd.copy_(self.ens.diagnostics(mid), non_blocking=True)
                ev = torch.cuda.Event()
                ev.record()
                retiring.append((ev, req, res, hs, hd))
                self.ens.remove(mid)
            while retiring and retiring[0][0].query():
                ev, req, res, hs, hd = retiring.pop(0)
                res.finished = time.perf_counter() - t0
                res.stats, res.diag = hs.tolist(), hd.tolist()
                results[req.rid] = res
            host["retire_s"] += time.perf_counter() - tr
        for ev, req, res, hs, hd in retiring:
            ev.synchronize()
            res.finished = time.perf_counter() - t0
            res.stats, res.diag = hs.tolist(), hd.tolist()
            results[req.rid] = res
        drain(0)
        t_end = time.perf_counter() - t0
        lats = np.array([x for r in results.values() for x in r.latencies_ms])
        span = t_end - (first_admit or 0.0)
        in_sys = np.array([r.finished - r.arrival for r in results.values()])
        metrics = {
            "requests": len(results),
            "member_steps": int(n_member_steps),
            "p50_step_ms": float(np.percentile(lats, 50)) if lats.size else math.nan,
            "p99_step_ms": float(np.percentile(lats, 99)) if lats.size else math.nan,
            "member_steps_per_s": n_member_steps / span if span > 0 else math.nan,
            "p50_time_in_system_s": float(np.percentile(in_sys, 50)) if in_sys.size else math.nan,
            "p99_time_in_system_s": float(np.percentile(in_sys, 99)) if in_sys.size else math.nan,
            "wall_s": t_end,
            "device_step_s": float(sum(ms for ms in step_ms)) / 1e3,
            "host_s": {k: round(v, 3) for k, v in host.items()},
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
    """The plain per-run path: one SLChannelFlow serving this request alone --
    its initial state (the same initialiser and dt as the server's admission)
    and its steps, one synchronisation at the end.  Returns (Result, seconds
    of the whole request).  The solver and its plans are built before the
    clock starts (a plain runner keeps them per grid, as the server keeps its
    initialisers)."""
    from .solver import SLChannelFlow

    flow = SLChannelFlow(config=req.config, device=device)
    _prepare(req, flow.device)  # per-grid initialiser and plans warm, as the server's
    torch.cuda.synchronize(flow.device)
    t = time.perf_counter()
    fields, dt = _prepare(req, flow.device)
    flow.set_fields(*fields)
    res = Result(rid=req.rid, dt=dt)
    torch.cuda.synchronize(flow.device)
    ts = time.perf_counter()
    for _ in range(req.n_steps):
        ub, fo = flow.step_sl_bdf2(dt)
    torch.cuda.synchronize(flow.device)
    el = time.perf_counter() - t
    res.steps = req.n_steps
    res.latencies_ms = [(time.perf_counter() - ts) * 1e3 / req.n_steps] * req.n_steps
    if keep_fields:
        res.fields = {"u": flow.u, "v": flow.v, "w": flow.w, "p": flow.p}
    res.stats = [ub.item(), fo.item(), float(flow.sl.n_clamped_last.item()), float(res.steps)]
    return res, el


def plain_fifo(requests, durations):
    """The plain per-run path under the same arrivals: requests served one at
    a time in arrival order, each taking its measured stand-alone duration.
    Returns member steps per second over the span and the p50 / p99 time in
    system."""
    t = 0.0
    first = None
    in_sys = []
    steps = 0
    for r, d in sorted(zip(requests, durations), key=lambda x: (x[0].arrival, x[0].rid)):
        start = max(t, r.arrival)
        first = start if first is None else first
        t = start + d
        in_sys.append(t - r.arrival)
        steps += r.n_steps
    span = t - (first or 0.0)
    return {"member_steps_per_s": steps / span if span > 0 else math.nan,
            "p50_time_in_system_s": float(np.percentile(in_sys, 50)),
            "p99_time_in_system_s": float(np.percentile(in_sys, 99))}
