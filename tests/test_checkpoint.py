"""Checkpoint files (slchannel_amd/checkpoint.py) in the reference's layout.

CPU: the keys and value types of utils.py:183-258 survive a write/read round
trip; a file written as the reference writes it (np.savez with those
keywords) loads; ``initialize_flow_from_file`` re-imposes the wall ghosts and
honours ``reset_time`` (initflow.py:225-286).

GPU: the restart semantics of tests/test_restart_sl.py of the reference --
10 steps, checkpoint, restart from the file through the config
(``initialization.field_file``), 10 more steps -- continue at step 10 and
stay within the reference's 1e-4 of the uninterrupted 20-step run (one BDF1
re-bootstrap, not bit-exact by design); and the restarted run equals the
oracle started from the same file (bootstrap, then BDF2) to 1e-12.
"""

import math
import os

import numpy as np
import pytest
import torch

from slchannel_amd.checkpoint import (KEYS, initialize_flow_from_file, load_flow_fields,
                                      save_flow_fields)


def _fields(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((nx + 1, ny + 2, nz + 2)),
            rng.standard_normal((nx + 2, ny + 1, nz + 2)),
            rng.standard_normal((nx + 2, ny + 2, nz + 1)),
            rng.standard_normal((nx + 2, ny + 2, nz + 2)))


def test_round_trip_layout(check, tmp_path):
    nx, ny, nz = 6, 5, 7
    u, v, w, p = _fields(nx, ny, nz, 1)
    zc, zf = np.linspace(-0.1, 2.1, nz + 2), np.linspace(0, 2, nz + 1)
    save_flow_fields(torch.tensor(u), torch.tensor(v), torch.tensor(w), torch.tensor(p),
                     torch.tensor(zc), torch.tensor(zf), 12.5, 6.25, 40, 0.4,
                     torch.tensor(0.0625), torch.tensor(-0.003), str(tmp_path), "f.npz")
    raw = np.load(str(tmp_path / "f.npz"))
    check("exactly the reference's keys", sorted(raw.files) == sorted(KEYS), f"{raw.files}")
    check("arrays float64, bit-identical",
          all(raw[k].dtype == np.float64 for k in ("u", "v", "w", "p", "z_c", "z_f"))
          and np.array_equal(raw["u"], u) and np.array_equal(raw["p"], p), "")
    check("scalars stored as 0-d", all(raw[k].shape == () for k in KEYS[6:]), "")
    f = load_flow_fields(str(tmp_path / "f.npz"))
    check("load: tensors and Python scalars",
          torch.equal(f["w"], torch.tensor(w)) and f["step"] == 40 and f["time"] == 0.4
          and f["u_tau"] == 0.0625 and f["forcing"] == -0.003 and f["Lx"] == 12.5, "")


def test_reference_written_file_restarts(check, tmp_path):
    nx, ny, nz = 6, 5, 7
    u, v, w, p = _fields(nx, ny, nz, 2)
    path = str(tmp_path / "fields.npz")
    # written the way utils.save_flow_fields does: np.savez, keyword per key
    np.savez(path, u=u, v=v, w=w, p=p, z_c=np.zeros(nz + 2), z_f=np.zeros(nz + 1), Lx=1.0,
             Ly=1.0, step=7, time=0.35, u_tau=0.05, forcing=0.001)
    uu, vv, ww, pp, step, t = initialize_flow_from_file(path)
    check("step and time continue", step == 7 and t == 0.35, f"{step} {t}")
    check("u, v no-slip ghost images",
          torch.equal(uu[:, :, 0], -uu[:, :, 1]) and torch.equal(uu[:, :, -1], -uu[:, :, -2])
          and torch.equal(vv[:, :, 0], -vv[:, :, 1]) and torch.equal(vv[:, :, -1], -vv[:, :, -2]),
          "")
    check("w zero on both wall faces",
          bool((ww[:, :, 0] == 0).all()) and bool((ww[:, :, -1] == 0).all()), "")
    check("interior untouched", np.array_equal(uu[:, :, 1:-1].numpy(), u[:, :, 1:-1])
          and np.array_equal(ww[:, :, 1:-1].numpy(), w[:, :, 1:-1]) and np.array_equal(pp.numpy(), p),
          "")
    _, _, _, _, step0, t0 = initialize_flow_from_file(path, reset_time=True)
    check("reset_time", step0 == 0 and t0 == 0.0, f"{step0} {t0}")
    missing = str(tmp_path / "none.npz")
    try:
        load_flow_fields(missing)
        ok = False
    except FileNotFoundError:
        ok = True
    check("missing file raises FileNotFoundError", ok, "")


DT = 0.01


def _cfg(extra=None):
    c = {"grid": {"nx": 16, "ny": 16, "nz": 32},
         "domain": {"Lx": 2 * math.pi, "Ly": 2 * math.pi, "Lz": 2.0},
         "flow": {"Re": 1000.0, "U_bulk": 1.0, "gamma": 1.5},
         "time": {"dt": DT}, "sl": {"interp_order": 6}}
    c.update(extra or {})
    return c


@pytest.mark.gpu
def test_restart_continues_and_matches(check, tmp_path):
    import oracle.oracle as orc
    from slchannel_amd import SLChannelFlow

    ref = SLChannelFlow(config=_cfg())
    ref.initialize(seed=5, urms=0.08, kmax=4.0)
    s1 = SLChannelFlow(config=_cfg())
    s1.set_fields(ref.u, ref.v, ref.w, ref.p)
    for _ in range(20):
        ref.step_sl_bdf2(DT)
    for _ in range(10):
        s1.step_sl_bdf2(DT)
    s1.save_checkpoint(str(tmp_path), "fields_ckpt.npz", u_tau=0.05)
    ckpt = str(tmp_path / "fields_ckpt.npz")
    s2 = SLChannelFlow(config=_cfg({"initialization": {"field_file": ckpt}}))
    check("restart continues from the checkpoint",
          s2.initial_step == 10 and abs(s2.initial_time - 10 * DT) < 1e-12
          and abs(s2.forcing - float(s1.forcing)) == 0.0,
          f"step={s2.initial_step} time={s2.initial_time} forcing={s2.forcing}")
    co = orc.ChannelOracle(16, 16, 32, 2 * math.pi, 2 * math.pi, 2.0, 1.5, 1000.0, order=6,
                           z_f=s2.z_f.numpy(), z_c=s2.z_c.numpy(), tanh_g=s2.plan.tanh_g)
    f = load_flow_fields(ckpt)
    co.set_fields(f["u"].numpy(), f["v"].numpy(), f["w"].numpy(), f["p"].numpy())
    worst = 0.0
    for n in range(10):
        s2.step_sl_bdf2(DT)
        if n < 3:
            co.step_sl_bdf2(DT)
            worst = max(worst, max(
                (a.cpu().numpy() - b).__abs__().max() / max(np.abs(b).max(), 1e-300)
                for a, b in ((s2.u, co.u), (s2.v, co.v), (s2.w, co.w))))
    check("restarted steps (bootstrap + BDF2) == oracle from the file", worst < 1e-12,
          f"max rel={worst:.1e}")
    umax = ref.u.abs().max().item()
    rel = max((a - b).abs().max().item() for a, b in ((ref.u, s2.u), (ref.v, s2.v),
                                                       (ref.w, s2.w))) / umax
    check("restart matches the uninterrupted run (reference bar 1e-4)", rel < 1e-4,
          f"rel_diff={rel:.2e} (bootstrap-limited, not bit-exact)")
    check("restart is a re-bootstrap, not a bit-exact continuation", rel > 0.0, f"{rel:.1e}")
