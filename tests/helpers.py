"""Shared test helpers: analytic staggered fields and backend adapters.

make_field builds a ghost-shaped array of one component with fn evaluated at
every entry's physical position, ghosts included, so interpolation tests need
no boundary machinery (the reference's tests/helpers.py:8-36 does the same).

Backends expose one interface over the CPU oracle ("oracle") and the HIP layer
("hip"); analytic tests are parametrised over both.
"""

import numpy as np


def positions(comp, nx, ny, dx, dy, z_f, z_c):
    if comp == "u":
        X = (np.arange(nx + 1) * dx)[:, None, None]
        Y = ((np.arange(ny + 2) - 0.5) * dy)[None, :, None]
        Z = np.asarray(z_c)[None, None, :]
    elif comp == "v":
        X = ((np.arange(nx + 2) - 0.5) * dx)[:, None, None]
        Y = (np.arange(ny + 1) * dy)[None, :, None]
        Z = np.asarray(z_c)[None, None, :]
    elif comp == "w":
        X = ((np.arange(nx + 2) - 0.5) * dx)[:, None, None]
        Y = ((np.arange(ny + 2) - 0.5) * dy)[None, :, None]
        Z = np.asarray(z_f)[None, None, :]
    else:
        raise ValueError(comp)
    return X, Y, Z


SHAPES = {
    "u": lambda nx, ny, nz: (nx + 1, ny + 2, nz + 2),
    "v": lambda nx, ny, nz: (nx + 2, ny + 1, nz + 2),
    "w": lambda nx, ny, nz: (nx + 2, ny + 2, nz + 1),
}


def make_field(comp, fn, nx, ny, nz, dx, dy, z_f, z_c):
    X, Y, Z = positions(comp, nx, ny, dx, dy, z_f, z_c)
    shape = SHAPES[comp](nx, ny, nz)
    return np.ascontiguousarray(np.broadcast_to(fn(X, Y, Z) * np.ones(shape), shape),
                                dtype=np.float64)


def arrival(comp, nx, ny, nz, dx, dy, z_f, z_c):
    """Arrival coordinates (broadcastable) of a component (semilag.py:292-307)."""
    ar = np.arange(1, nx + 1, dtype=np.float64)
    aj = np.arange(1, ny + 1, dtype=np.float64)
    if comp == "u":
        return (ar * dx)[:, None, None], ((aj - 0.5) * dy)[None, :, None], \
            np.asarray(z_c)[1 : nz + 1][None, None, :]
    if comp == "v":
        return ((ar - 0.5) * dx)[:, None, None], (aj * dy)[None, :, None], \
            np.asarray(z_c)[1 : nz + 1][None, None, :]
    return ((ar - 0.5) * dx)[:, None, None], ((aj - 0.5) * dy)[None, :, None], \
        np.asarray(z_f)[1:nz][None, None, :]


class OracleBackend:
    name = "oracle"

    def __init__(self, nx, ny, nz, Lx, Ly, Lz, z_f, z_c, gamma, order=4, n_traj_iters=2,
                 stretching="symmetric"):
        from oracle import oracle as orc

        self.sl = orc.SLOracle(nx, ny, nz, Lx / nx, Ly / ny, Lx, Ly, Lz, z_f, z_c, gamma,
                               stretching_type=stretching, order=order,
                               n_traj_iters=n_traj_iters)
        self.shape = lambda c: self.sl.arrival_shape(c)

    def departure(self, comp, um, vm, wm, dt):
        return tuple(a.reshape(self.shape(comp)) for a in self.sl.departure(comp, um, vm, wm, dt))

    def advect(self, u, v, w, um, vm, wm, dt):
        out = self.sl.advect(u, v, w, um, vm, wm, dt)
        return out, self.sl.n_clamped_last


class HipBackend:
    name = "hip"

    def __init__(self, nx, ny, nz, Lx, Ly, Lz, z_f, z_c, gamma, order=4, n_traj_iters=2,
                 stretching="symmetric"):
        from slchannel_amd import SLAdvector

        self.adv = SLAdvector(nx, ny, nz, Lx / nx, Ly / ny, Lx, Ly, Lz, z_f, z_c, gamma,
                              stretching_type=stretching, order=order,
                              n_traj_iters=n_traj_iters, device="cuda")

    @staticmethod
    def _d(a):
        import torch

        return torch.as_tensor(np.ascontiguousarray(a)).cuda()

    def departure(self, comp, um, vm, wm, dt):
        feet = self.adv.departure_coords(comp, dt, self._d(um), self._d(vm), self._d(wm))
        return tuple(t.cpu().numpy() for t in feet)

    def advect(self, u, v, w, um, vm, wm, dt):
        d = self._d
        out = self.adv.advect(d(u), d(v), d(w), d(um), d(vm), d(wm), dt)
        return tuple(t.cpu().numpy() for t in out), int(self.adv.n_clamped_last.item())


def backend(name, *a, **k):
    return (OracleBackend if name == "oracle" else HipBackend)(*a, **k)
