"""Wall-normal grid of the channel (host side, float64).

generate_grid restates slchannel/utils.py:21-61 with torch on the CPU so the
tables handed to the HIP plan are the numbers the reference computes on the
same machine: tanh-stretched faces z_f (nz+1), face-midpoint centres z_c with
mirrored ghost centres (nz+2), and the spacings dz_f (nz), dz_c (nz+1).
"""

import torch


def generate_grid(gamma, nz, Lz, device="cpu", stretching_type="symmetric"):
    """utils.py:21-61: returns (z_f, z_c, dz_f, dz_c) as float64 tensors."""
    k = torch.linspace(0, nz, nz + 1, dtype=torch.float64)
    g = torch.tensor(gamma, dtype=torch.float64)
    if stretching_type == "bottom":
        xi = k / nz
        z_f = Lz * (1.0 - torch.tanh(gamma * (1.0 - xi)) / torch.tanh(g))
    elif stretching_type == "symmetric":
        xi = (2 * k / nz) - 1
        z_f = 0.5 * Lz * (1 + torch.tanh(gamma * xi) / torch.tanh(g))
    else:
        raise ValueError(f"stretching_type must be 'symmetric' or 'bottom', got {stretching_type!r}")
    z_c_inn = 0.5 * (z_f[:-1] + z_f[1:])
    z_c = torch.cat([-z_c_inn[:1], z_c_inn, (2 * z_f[-1] - z_c_inn[-1]).reshape(1)])
    dz_f = z_f[1:] - z_f[:-1]
    dz_c = z_c[1:] - z_c[:-1]
    return tuple(t.to(device) for t in (z_f, z_c, dz_f, dz_c))
