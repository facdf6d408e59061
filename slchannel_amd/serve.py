"""Serving loop: continuous batching of independent channel runs on one GPU.

Requests (a channel config, a seed, a number of steps, an arrival time)
arrive over time.  The loop initialises admitted requests on the device
(slc_init_solenoidal, a few launches per request) and advances the admitted
runs in two lanes that step concurrently on their own streams:

* the solo lane: one large request (more than `small_points` grid points,
  one step fills the device) at a time, in arrival order, stepped by the
  single-run solver (SLChannelFlow, the plain per-run path's kernels, so it
  runs at the plain path's speed and gives its doubles);
* the batch: small requests advanced together by the batched ensemble (one
  launch per substep for all of them, include/slc_ens.h).  The oldest
  waiting small request joins when the batch is empty or its grid points fit
  the batch budget (`batch_points`); any arrived small request joins while
  the small runs' points fit the budget -- it rides in the launches of the
  batch, costing its own arithmetic and no launches of its own, beside the
  solo run's steps.

Admission needs the server's HBM footprint (ensemble, initialisers, solo
solvers) plus the newcomer's (slc_ens_member_bytes) to stay within the
budget.  Runs join and leave between steps: a run admitted while others are
mid-flight bootstraps with BDF1 inside the same launches in which the others
take BDF2 steps, and a run that has done its steps is copied out and its
memory released, so a waiting request can take its place.

Reported per run: arrival, admission and completion times, every step's
latency (the device time of the step it took part in, from hipEvents), its
final run diagnostics (max|div|, u_tau, CFL, non-finite entries;
slc_ens_diagnostics / slc_diagnostics), a status, and optionally its final
fields.
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
from .ensemble import (N_STATS, OPT_POOL, Ensemble, _INIT_CACHE, _config_grid, initial_fields,
                       member_bytes)


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


# Per grid and device: the single-run solver that steps the server's large
# requests alone (the plain per-run path), apart from the initialiser of
# the same grid so that the next request's initial state can be built while
# a run steps.
_SOLO_CACHE = {}


def _solo_key(config, device):
    """The solo solver's cache key: the whole config except the bulk Reynolds
    number and velocity, which are set per run (set_reynolds, U_bulk)."""
    import json

    cfg = json.loads(json.dumps(config))
    cfg.get("flow", {}).pop("Re", None)
    cfg.get("flow", {}).pop("U_bulk", None)
    return (json.dumps(cfg, sort_keys=True), str(device))


def _solo_flow(config, device):
    """The solo solver for this request's grid and options, set to its Re_b
    and U_b (built once per key)."""
    from .solver import SLChannelFlow

    key = _solo_key(config, device)
    flow = _SOLO_CACHE.get(key)
    if flow is None:
        flow = _SOLO_CACHE[key] = SLChannelFlow(config=config, device=device)
    flow.set_reynolds(float(config["flow"]["Re"]))
    flow.U_bulk = float(config["flow"].get("U_bulk", 1.0))
    return flow


def init_cache_bytes(device):
    """Device bytes held by the per-grid initialisers and solo solvers of
    `device` (their plans, fields and histories), which share the HBM budget
    with the members."""
    def tensors(obj):
        out = []
        for t in vars(obj).values():
            if isinstance(t, torch.Tensor) and t.is_cuda:
                out.append(t)
            elif isinstance(t, (list, tuple)):
                out.extend(x for x in t if isinstance(x, torch.Tensor) and x.is_cuda)
        return out

    total = 0
    for key, flow in list(_INIT_CACHE.items()) + list(_SOLO_CACHE.items()):
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
        # released member arenas stay pooled for reuse up to the member cap:
        # the budget check counts them and trims the oldest when a request
        # needs their room, so a retirement never waits on hipFree
        self.ens.set_option(OPT_POOL, int(max_members))
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
        # stay ordered; the solo lane steps on a stream of its own (default
        # priority, as every stream here)
        self._streams = {}
        self._solo_stream = torch.cuda.Stream(self.device)

    def _stream_for(self, config):
        key = _grid_key(config)
        st = self._streams.get(key)
        if st is None:
            st = self._streams[key] = torch.cuda.Stream(self.device)
        return st

    def warmup(self, configs):
        """Server start-up for a grid catalogue: builds the per-grid initialiser
        (its rocFFT plans) of every distinct grid, and the solo solver of every
        large one, before requests arrive."""
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
            if _points(cfg) > self.small_points:
                _solo_flow(cfg, self.device)
        torch.cuda.synchronize(self.device)

    def footprint(self):
        """HBM the server holds: the ensemble (live members, pooled arenas,
        class buffers) plus the per-grid initialisers."""
        return self.ens.bytes() + init_cache_bytes(self.device)

    def fits(self, req, reserved=0, preparing=0, trim=False, solo=False):
        """Whether the request fits the budget beside what the server holds.
        Released arenas kept for reuse count as held; with trim=True, the
        oldest of them are freed when that makes the request fit.  A solo run
        (solo=True) needs nothing more once its grid's solver exists (the
        solver is counted in the footprint), else about a member's bytes."""
        if solo:
            if _solo_key(req.config, self.device) in _SOLO_CACHE:
                return True
        elif len(self.ens) + preparing >= self.max_members:
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

    def _pick(self, pending, now, active_pts, small_pts, reserved, preparing, solo_busy):
        """The next request to admit and its lane ("solo" or "ens"), or None.
        The oldest arrived request goes first: a large one (more than
        `small_points` grid points: one step fills the device) to the solo
        lane once it is free, a small one into the batch when the batch is
        empty or it fits the batch budget.  Otherwise the oldest arrived small
        request that fits the small runs' budget joins the batch.  All need
        the HBM budget."""
        if not pending or pending[0].arrival > now:
            return None
        head = _points(pending[0].config)
        if head > self.small_points:
            if not solo_busy and self.fits(pending[0], reserved, preparing, trim=True,
                                           solo=True):
                return 0, "solo"
        elif (active_pts == 0 or active_pts + head <= self.batch_points) and \
                self.fits(pending[0], reserved, preparing, trim=True):
            return 0, "ens"
        for i, r in enumerate(pending[1:], 1):
            if r.arrival > now:
                break
            p = _points(r.config)
            if p <= self.small_points and small_pts + p <= self.batch_points and \
                    self.fits(r, reserved, preparing, trim=True):
                return i, "ens"
        return None

    def run(self, requests, log=None, inflight=3):
        """Serve `requests` (arrival times relative to the start); returns
        (results by rid, metrics dict).  Steps are pipelined: the host keeps
        at most `inflight` steps of each lane queued ahead of the device and
        only synchronises to retire a finished run; each step's latency is its
        device time from hipEvents.  The solo lane (one large run, the
        single-run solver on its own stream) and the batch run concurrently."""
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
        queued = []  # batch steps: (start event, end event, results stepped)
        solo = None  # the solo run: dict(req, res, flow, dt)
        solo_q = []  # solo steps: (start event, end event, result)
        retiring = []  # (event after the copies out, request, result, pinned stats, diag)
        preparing = []  # (event, request, fields, dt, pinned probe, stream, bytes, lane)
        guards = {}  # member id or "solo" -> BlowupGuard
        checks = []  # (event, member id or "solo", rid, pinned diagnostics)
        t0 = time.perf_counter()
        n_member_steps = 0
        first_admit = None
        host = {"admit_s": 0.0, "step_s": 0.0, "retire_s": 0.0, "wait_s": 0.0}

        step_ms = []
        pts = {}  # member id -> grid points

        def drain(q, keep):
            while len(q) > keep:
                ev0, ev1, stepped = q.pop(0)
                ev1.synchronize()
                ms = ev0.elapsed_time(ev1)
                step_ms.append(ms)
                for res in stepped:
                    res.latencies_ms.append(ms)

        def copy_out(res, stats, diag, stream):
            """Pinned copies of a finished run's stats and diagnostics, read
            once the returned event has completed."""
            hs = torch.empty(N_STATS, dtype=torch.float64, pin_memory=True)
            hd = torch.empty(5, dtype=torch.float64, pin_memory=True)
            with torch.cuda.stream(stream):
                hs.copy_(stats, non_blocking=True)
                hd.copy_(diag, non_blocking=True)
                ev = torch.cuda.Event()
                ev.record()
            return ev, hs, hd

        def retire_solo(status):
            nonlocal solo
            req, res, flow, dt = solo["req"], solo["res"], solo["flow"], solo["dt"]
            res.status = status
            st = self._solo_stream
            with torch.cuda.stream(st):
                if self.keep_fields:
                    res.fields = {k: getattr(flow, k).clone() for k in "uvwp"}
                diag = torch.empty(5, dtype=torch.float64, device=self.device)
                flow.plan.diagnostics(flow.u, flow.v, flow.w, flow.p, dt, diag)
                b = flow._bulk
                stats = torch.stack([b[2], b[1], flow.sl.n_clamped_last.to(torch.float64),
                                     torch.full_like(b[0], float(res.steps)), b[3]])
            ev, hs, hd = copy_out(res, stats, diag, st)
            retiring.append((ev, req, res, hs, hd))
            guards.pop("solo", None)
            solo = None

        def n_running():
            return len(active) + (1 if solo is not None else 0)

        while pending or active or preparing or solo is not None:
            now = time.perf_counter() - t0
            ta = time.perf_counter()
            # admission, FIFO over the requests that have arrived: the initial
            # state is launched on the grid's stream and read back (dt, the
            # finiteness check) through pinned memory once that stream is done
            reserved = sum(x[6] for x in preparing if x[7] == "ens")
            ens_prep = [x for x in preparing if x[7] == "ens"]
            active_pts = sum(pts.values()) + sum(_points(x[1].config) for x in ens_prep)
            small_pts = sum(p for p in pts.values() if p <= self.small_points) + sum(
                _points(x[1].config) for x in ens_prep
                if _points(x[1].config) <= self.small_points)
            solo_busy = solo is not None or any(x[7] == "solo" for x in preparing)
            pick = self._pick(pending, now, active_pts, small_pts, reserved, len(ens_prep),
                              solo_busy)
            if pick is not None:
                req = pending.pop(pick[0])
                st = self._stream_for(req.config)
                with torch.cuda.stream(st):
                    fields, dt, probe = _prepare_async(req, self.device)
                    h = torch.empty(2, dtype=torch.float64, pin_memory=True)
                    h.copy_(probe, non_blocking=True)
                    ev = torch.cuda.Event()
                    ev.record()
                preparing.append((ev, req, fields, dt, h, st,
                                  member_bytes(req.config, self.storage), pick[1]))
            while preparing and preparing[0][0].query():
                ev, req, fields, dt, h, st, _, lane = preparing.pop(0)
                probe = h.tolist()
                if dt is None:
                    dt = _dt_of(req, probe[0])
                if probe[1] > 0:
                    res = Result(rid=req.rid, arrival=req.arrival, dt=dt,
                                 status="rejected_nonfinite")
                    res.admitted = res.finished = time.perf_counter() - t0
                    results[req.rid] = res
                    continue
                res = Result(rid=req.rid, dt=dt)
                res.arrival = req.arrival
                if lane == "solo":
                    flow = _solo_flow(req.config, self.device)
                    sst = self._solo_stream
                    sst.wait_event(ev)
                    for t in fields:
                        if t is not None:
                            t.record_stream(sst)
                    with torch.cuda.stream(sst):
                        flow.set_fields(*fields)
                    solo = {"req": req, "res": res, "flow": flow, "dt": dt}
                    guards["solo"] = BlowupGuard(req.u_tau_nominal, self.blowup_factor)
                else:
                    with torch.cuda.stream(st):
                        mid = self.ens.add(req.config, dt, fields=fields)
                    guards[mid] = BlowupGuard(req.u_tau_nominal, self.blowup_factor)
                    active[mid] = (req, res)
                    pts[mid] = _points(req.config)
                res.admitted = time.perf_counter() - t0
                first_admit = res.admitted if first_admit is None else first_admit
                if log:
                    c = _config_grid(req.config)
                    log(f"admit {req.rid} {c['nx']}x{c['ny']}x{c['nz']} "
                        f"steps={req.n_steps} footprint={self.footprint()} "
                        f"active={n_running()} lane={lane}")
            host["admit_s"] += time.perf_counter() - ta
            if not active and solo is None and not preparing and pending and \
                    pending[0].arrival <= now and not self.fits(
                        pending[0], solo=_points(pending[0].config) > self.small_points):
                # nothing runs that could free memory: the server's fixed
                # footprint (pooled arenas, initialisers) leaves no room
                req = pending.pop(0)
                res = Result(rid=req.rid, arrival=req.arrival, status="rejected_invalid")
                res.admitted = res.finished = time.perf_counter() - t0
                res.diag = ["does not fit the HBM budget beside the server's fixed footprint"]
                results[req.rid] = res
                continue
            if not active and solo is None:
                if preparing:
                    time.sleep(1e-4)
                elif pending:
                    wait = pending[0].arrival - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.05))
                continue
            tw = time.perf_counter()
            if solo is not None:
                drain(solo_q, inflight - 1)
            if active:
                drain(queued, inflight - 1)
            ts = time.perf_counter()
            host["wait_s"] += ts - tw
            done = []
            if solo is not None:
                flow, res, req = solo["flow"], solo["res"], solo["req"]
                with torch.cuda.stream(self._solo_stream):
                    ev0 = torch.cuda.Event(enable_timing=True)
                    ev1 = torch.cuda.Event(enable_timing=True)
                    ev0.record()
                    flow.step_sl_bdf2(solo["dt"])
                    ev1.record()
                solo_q.append((ev0, ev1, [res]))
                n_member_steps += 1
                res.steps += 1
                if res.steps >= req.n_steps:
                    done.append(("solo", "ok"))
                elif self.guard_every and res.steps % self.guard_every == 0:
                    with torch.cuda.stream(self._solo_stream):
                        d = torch.empty(5, dtype=torch.float64, device=self.device)
                        flow.plan.diagnostics(flow.u, flow.v, flow.w, flow.p, solo["dt"], d)
                        hg = torch.empty(5, dtype=torch.float64, pin_memory=True)
                        hg.copy_(d, non_blocking=True)
                        ev = torch.cuda.Event()
                        ev.record()
                    checks.append((ev, "solo", req.rid, hg))
            if active:
                ev0 = torch.cuda.Event(enable_timing=True)
                ev1 = torch.cuda.Event(enable_timing=True)
                ev0.record()
                self.ens.step()
                ev1.record()
                queued.append((ev0, ev1, [res for _, res in active.values()]))
                n_member_steps += len(active)
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
            host["step_s"] += time.perf_counter() - ts
            # guard results whose copies have landed (no waiting)
            while checks and checks[0][0].query():
                _, mid, rid, h = checks.pop(0)
                running = (solo is not None and solo["req"].rid == rid) if mid == "solo" \
                    else (mid in active and active[mid][0].rid == rid)
                if running and all(m != mid for m, _ in done):
                    verdict = guards[mid].check(h.tolist())
                    if verdict is not None:
                        done.append((mid, "stopped_" + verdict))
            tr = time.perf_counter()
            for mid, status in done:
                if mid == "solo":
                    retire_solo(status)
                    continue
                req, res = active.pop(mid)
                guards.pop(mid, None)
                pts.pop(mid, None)
                res.status = status
                # copies are stream-ordered behind the member's last step;
                # the arena is reused only by work queued after them
                if self.keep_fields:
                    res.fields = self.ens.get_fields(mid)
                ev, hs, hd = copy_out(res, self.ens.stats(mid), self.ens.diagnostics(mid),
                                      torch.cuda.current_stream(self.device))
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
        drain(queued, 0)
        drain(solo_q, 0)
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
