"""Checkpoint files in the reference's npz layout.

``save_flow_fields`` / ``load_flow_fields`` keep the keys and value types of
slchannel/utils.py:183-258 (u, v, w, p, z_c, z_f, Lx, Ly, step, time, u_tau,
forcing), so a ``fields.npz`` written by either code restarts the other.
``initialize_flow_from_file`` follows slchannel/initflow.py:225-286: fields,
step and time from the file (or zero with ``reset_time``) and the wall ghost
values re-imposed.  A restarted SL-BDF2 run takes one BDF1 bootstrap step
first -- the histories are not in the file (solver.py:834-844).

File I/O and a few slice assignments only; the arrays go straight to the
requested device.
"""

import os

import numpy as np
import torch

KEYS = ("u", "v", "w", "p", "z_c", "z_f", "Lx", "Ly", "step", "time", "u_tau", "forcing")


def _host(a):
    return a.detach().cpu().numpy() if torch.is_tensor(a) else a


def _scalar(a):
    return a.item() if torch.is_tensor(a) else a


def save_flow_fields(u, v, w, p, z_c, z_f, Lx, Ly, step, time, u_tau, forcing, results_folder,
                     filename="fields.npz"):
    """utils.py:183-215: one uncompressed npz, overwritten in place."""
    np.savez(os.path.join(results_folder, filename), u=_host(u), v=_host(v), w=_host(w),
             p=_host(p), z_c=_host(z_c), z_f=_host(z_f), Lx=Lx, Ly=Ly, step=step, time=time,
             u_tau=_scalar(u_tau), forcing=_scalar(forcing))


def load_flow_fields(filepath, device="cpu"):
    """utils.py:217-258: tensors on ``device`` for the arrays, Python
    numbers for the scalars."""
    if not os.path.exists(filepath):
        raise FileNotFoundError(f"Field file not found: {filepath}")
    data = np.load(filepath)  # allow_pickle stays False: arrays and scalars only
    return {
        "u": torch.tensor(data["u"], device=device),
        "v": torch.tensor(data["v"], device=device),
        "w": torch.tensor(data["w"], device=device),
        "p": torch.tensor(data["p"], device=device),
        "z_c": torch.tensor(data["z_c"], device=device),
        "z_f": torch.tensor(data["z_f"], device=device),
        "Lx": float(data["Lx"]),
        "Ly": float(data["Ly"]),
        "step": int(data["step"]),
        "time": float(data["time"]),
        "u_tau": float(data["u_tau"]),
        "forcing": float(data["forcing"]),
    }


def initialize_flow_from_file(field_file, device="cpu", reset_time=False):
    """initflow.py:225-286: (u, v, w, p, initial_step, initial_time) with the
    wall ghosts re-imposed as the reference does (no-slip images for u and v
    at both ends, w = 0 on both wall faces)."""
    f = load_flow_fields(field_file, device=device)
    u, v, w, p = f["u"], f["v"], f["w"], f["p"]
    if reset_time:
        initial_step, initial_time = 0, 0.0
    else:
        initial_step, initial_time = int(f["step"]), float(f["time"])
    u[:, :, 0] = -u[:, :, 1]
    u[:, :, -1] = -u[:, :, -2]
    v[:, :, 0] = -v[:, :, 1]
    v[:, :, -1] = -v[:, :, -2]
    w[:, :, 0] = 0.0
    w[:, :, -1] = 0.0
    return u, v, w, p, initial_step, initial_time
