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


def _interp1d_weights(coords, nodes):
    """initflow.py:20-28: lower source index and upper-node weight of each
    coordinate in the sorted nodes (searchsorted left, clamped into
    [1, len-1]; weights clamped to [0, 1], i.e. constant extrapolation)."""
    k = np.clip(np.searchsorted(nodes, coords, side="left"), 1, len(nodes) - 1)
    k0 = k - 1
    denom = np.maximum(nodes[k] - nodes[k0], 1e-300)
    t = np.clip((coords - nodes[k0]) / denom, 0.0, 1.0)
    return k0.astype(np.int32), t


def _trilinear_device(src, xn, yn, zn, xq, yq, zq, device):
    """initflow.py:31-44 on the device (slc_interp_trilinear): the 1-D
    tables here, the tensor-product sums in one kernel."""
    import ctypes

    from . import _lib

    lib = _lib.load()
    tabs = [_interp1d_weights(np.asarray(q, dtype=np.float64), np.asarray(n, dtype=np.float64))
            for q, n in ((xq, xn), (yq, yn), (zq, zn))]
    dev = [(torch.as_tensor(k, device=device), torch.as_tensor(t, device=device))
           for k, t in tabs]
    s = torch.as_tensor(np.ascontiguousarray(src), dtype=torch.float64, device=device)
    out = torch.empty((len(xq), len(yq), len(zq)), dtype=torch.float64, device=device)
    ptr = lambda a: ctypes.c_void_p(a.data_ptr())  # noqa: E731
    stream = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    _lib.check(lib.slc_interp_trilinear(
        device.index or 0, ptr(s), s.shape[0], s.shape[1], s.shape[2],
        ptr(dev[0][0]), ptr(dev[0][1]), len(xq), ptr(dev[1][0]), ptr(dev[1][1]), len(yq),
        ptr(dev[2][0]), ptr(dev[2][1]), len(zq), ptr(out), stream), "slc_interp_trilinear")
    return out


def interpolated_fields(field_file, nx, ny, nz, Lx, Ly, Lz, z_c, z_f, device,
                        source_half="lower"):
    """initialize_flow_interpolated (initflow.py:47-120): a stored field,
    possibly of another grid and box, onto this staggered grid.  The source
    grid comes from the file (z tables, Lx, Ly, array shapes; x/y uniform);
    its periodic ghosts are refreshed; x and y map proportionally onto the
    source box (a wider or longer source is stretched onto the target), z
    directly (a taller source gives its lower part, source_half 'lower');
    each component is interpolated trilinearly at its own staggered nodes on
    the device.  Returns u, v, w and a zero p on `device`; the caller fills
    the ghosts, rescales u to the bulk velocity and projects (solver.py:351-401,
    599-618)."""
    device = torch.device(device)
    src = load_flow_fields(field_file, device="cpu")
    u_s, v_s, w_s = (src[k].numpy().copy() for k in "uvw")
    z_c_s, z_f_s = src["z_c"].numpy(), src["z_f"].numpy()
    Lx_s, Ly_s = float(src["Lx"]), float(src["Ly"])
    nxs, nys = u_s.shape[0] - 1, v_s.shape[1] - 1
    Lz_s = float(z_f_s[-1])
    if source_half != "lower" and Lz_s > Lz:
        raise ValueError(f"source_half='{source_half}' not supported (only 'lower')")
    dxs, dys = Lx_s / nxs, Ly_s / nys
    # periodic ghost layers of the source (initflow.py:83-89)
    u_s[0, :, :] = u_s[-1, :, :]
    u_s[:, 0, :] = u_s[:, -2, :]
    u_s[:, -1, :] = u_s[:, 1, :]
    v_s[0, :, :] = v_s[-2, :, :]
    v_s[-1, :, :] = v_s[1, :, :]
    v_s[:, 0, :] = v_s[:, -1, :]
    w_s[0, :, :] = w_s[-2, :, :]
    w_s[-1, :, :] = w_s[1, :, :]
    w_s[:, 0, :] = w_s[:, -2, :]
    w_s[:, -1, :] = w_s[:, 1, :]
    # source and target node coordinates (initflow.py:94-112)
    x_face_s = np.arange(nxs + 1, dtype=np.float64) * dxs
    x_cent_s = (np.arange(nxs + 2, dtype=np.float64) - 0.5) * dxs
    y_face_s = np.arange(nys + 1, dtype=np.float64) * dys
    y_cent_s = (np.arange(nys + 2, dtype=np.float64) - 0.5) * dys
    dx, dy = Lx / nx, Ly / ny
    rx, ry = Lx_s / Lx, Ly_s / Ly
    x_face_t = np.arange(nx + 1, dtype=np.float64) * dx * rx
    x_cent_t = np.clip((np.arange(nx + 2, dtype=np.float64) - 0.5) * dx * rx,
                       x_cent_s[0], x_cent_s[-1])
    y_face_t = np.arange(ny + 1, dtype=np.float64) * dy * ry
    y_cent_t = np.clip((np.arange(ny + 2, dtype=np.float64) - 0.5) * dy * ry,
                       y_cent_s[0], y_cent_s[-1])
    z_c, z_f = _host(z_c), _host(z_f)
    u = _trilinear_device(u_s, x_face_s, y_cent_s, z_c_s, x_face_t, y_cent_t, z_c, device)
    v = _trilinear_device(v_s, x_cent_s, y_face_s, z_c_s, x_cent_t, y_face_t, z_c, device)
    w = _trilinear_device(w_s, x_cent_s, y_cent_s, z_f_s, x_cent_t, y_cent_t, z_f, device)
    p = torch.zeros((nx + 2, ny + 2, nz + 2), dtype=torch.float64, device=device)
    return u, v, w, p
