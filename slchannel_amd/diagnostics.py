"""Run diagnostics: dt sizing (torch ops on the device, off the hot path)
and the blow-up guard over the device diagnostics of slc_diagnostics.

compute_cfl restates compute_cfl_fused (slchannel/operators.py:839-940): the
inverse trajectory time-scale max over cells of (|u|/dx + |v|/dy + |w|/dz)
with the staggered components averaged to each face family.  The SL scheme's
dt is CFL_target / compute_cfl (solver.py:652-674, without the dt clamps).
"""

import torch


def compute_cfl(u, v, w, nx, ny, nz, dx, dy, dz_f, dz_c):
    """Returns a float (one host synchronisation)."""
    dz_f = torch.as_tensor(dz_f, dtype=u.dtype, device=u.device)
    dz_c = torch.as_tensor(dz_c, dtype=u.dtype, device=u.device)
    dxi, dyi = 1.0 / dx, 1.0 / dy
    dzfi, dzci = 1.0 / dz_f, 1.0 / dz_c
    I, J = slice(1, nx + 1), slice(1, ny + 1)
    K = slice(1, nz + 1)
    ux = u[I, J, K].abs()
    vx = 0.25 * (v[I, J, K] + v[I, 0:ny, K] + v[2:nx + 2, J, K] + v[2:nx + 2, 0:ny, K]).abs()
    wx = 0.25 * (w[I, J, K] + w[I, J, 0:nz] + w[2:nx + 2, J, K] + w[2:nx + 2, J, 0:nz]).abs()
    dtix = ux * dxi + vx * dyi + wx * dzfi[0:nz].view(1, 1, -1)
    uy = 0.25 * (u[I, J, K] + u[I, 2:ny + 2, K] + u[0:nx, 2:ny + 2, K] + u[0:nx, J, K]).abs()
    vy = v[I, J, K].abs()
    wy = 0.25 * (w[I, J, K] + w[I, 2:ny + 2, K] + w[I, 2:ny + 2, 0:nz] + w[I, J, 0:nz]).abs()
    dtiy = uy * dxi + vy * dyi + wy * dzfi[0:nz].view(1, 1, -1)
    Kz, Kz1 = slice(1, nz), slice(2, nz + 1)
    uz = 0.25 * (u[I, J, Kz] + u[0:nx, J, Kz] + u[0:nx, J, Kz1] + u[I, J, Kz1]).abs()
    vz = 0.25 * (v[I, J, Kz] + v[I, 0:ny, Kz] + v[I, 0:ny, Kz1] + v[I, J, Kz1]).abs()
    wz = w[I, J, Kz].abs()
    dtiz = uz * dxi + vz * dyi + wz * dzci[1:nz].view(1, 1, -1)
    return float(torch.stack([dtix.max(), dtiy.max(), dtiz.max()]).max().item())


def cfl_dt(flow, cfl_target):
    """dt giving the requested trajectory CFL on the flow's current fields."""
    dti = compute_cfl(flow.u, flow.v, flow.w, flow.nx, flow.ny, flow.nz, flow.dx, flow.dy,
                      flow.dz_f, flow.dz_c)
    return cfl_target / (dti if dti > 1e-10 else 1.0)


class BlowupGuard:
    """solver.py:1202-1213 (factor: time.blowup_u_tau_factor, default 2,
    solver.py:234): True once u_tau has exceeded factor x nominal on
    three consecutive diagnostics (a single spike resets).  Fed from rows of
    SLChannelFlow.diagnostics() whenever the host reads them; a row with
    non-finite entries (its [4] > 0) trips the NaN/Inf guard at once
    (solver.py:1436-1479)."""

    def __init__(self, u_tau_nominal, factor=2.0, strikes=3):
        self.u_tau_nominal = u_tau_nominal
        self.factor = float(factor)
        self.strikes_needed = int(strikes)
        self.strikes = 0

    def check(self, row):
        """row: sequence [max_div, u_tau, cfl, dti, nonfinite].  Returns None,
        'nonfinite' or 'blowup'."""
        if float(row[4]) > 0:
            return "nonfinite"
        if self.u_tau_nominal is not None and float(row[1]) > self.factor * self.u_tau_nominal:
            self.strikes += 1
        else:
            self.strikes = 0
        return "blowup" if self.strikes >= self.strikes_needed else None
