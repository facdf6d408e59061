"""Seeded solenoidal initial fields for the channel (host side, numpy).

The reference has no seeded divergence-free random initialiser
(slchannel/initflow.py:123-223 offers parabolic / vortices / uniform), so this
one is written here.  Base profile u = 1.5 U_b (1 - zeta^2), zeta = 2z/Lz - 1
(the 'parabolic' profile of initflow.py:140-145, bulk U_b), plus the curl of a
random vector potential

    A_c(x, y, z) = sum_{m,n,l} Re[a^c_{mnl} e^{i(k_m x + k_n y)}] G_l(z),
    G_l(z) = (1 - zeta^2)^2 sin(l pi z / Lz),

so each mode is exactly solenoidal and u, v, w and their wall-normal
derivatives of A vanish at both walls.  Modes are integer triples with
kmin <= |(m, n, l)| <= kmax; amplitudes a ~ N(0,1) + i N(0,1) scaled by
|k|^-p (p = 1: flat per-mode velocity spectrum; p = 17/6: E(k) ~ k^-5/3),
drawn from numpy's PCG64 with the given seed.  The perturbation is scaled to
the requested u'rms = sqrt(<u'^2 + v'^2 + w'^2>/3) over the interiors.
Velocities are evaluated analytically at every staggered interior position;
ghosts are left at zero for the caller's boundary fill.  The discrete field
is then projected (the caller runs the projection with dt = 1, as
solver.py:599-618 does for a fresh or restarted field).
"""

import math

import numpy as np


def _modes(kmin, kmax, nx, ny):
    K = int(math.floor(kmax))
    K = min(K, nx // 2 - 1, ny // 2 - 1)
    out = []
    for m in range(-K, K + 1):
        for n in range(-K, K + 1):
            for l in range(1, int(math.floor(kmax)) + 1):
                k = math.sqrt(m * m + n * n + l * l)
                if kmin <= k <= kmax:
                    out.append((m, n, l, k))
    return out


def solenoidal_fields(nx, ny, nz, Lx, Ly, Lz, z_f, z_c, seed, urms=0.1, kmin=1.0, kmax=8.0,
                      spectrum="flat", U_bulk=1.0, chunk=32, device=None):
    """Return ghost-shaped arrays (u, v, w) with interiors set: numpy arrays,
    or float64 torch tensors on `device` when one is given (the x-y
    synthesis then runs as torch FFTs there; same modes and amplitudes, FFT
    rounding aside)."""
    z_f = np.asarray(z_f, dtype=np.float64)
    z_c = np.asarray(z_c, dtype=np.float64)
    rng = np.random.Generator(np.random.PCG64(seed))
    modes = _modes(kmin, kmax, nx, ny)
    if not modes:
        raise ValueError("no Fourier modes in the requested band")
    p = {"flat": 1.0, "k-5/3": 17.0 / 6.0}[spectrum]
    M = len(modes)
    mm = np.array([t[0] for t in modes])
    nn = np.array([t[1] for t in modes])
    ll = np.array([t[2] for t in modes])
    kk = np.array([t[3] for t in modes])
    amp = kk ** (-p)
    A = {c: (rng.standard_normal(M) + 1j * rng.standard_normal(M)) * amp for c in "xyz"}
    kx = 2 * math.pi * mm / Lx
    ky = 2 * math.pi * nn / Ly
    dx, dy = Lx / nx, Ly / ny

    def G(z):
        zeta = 2.0 * z / Lz - 1.0
        env = (1.0 - zeta**2) ** 2
        denv = -8.0 * zeta * (1.0 - zeta**2) / Lz
        arg = np.outer(ll, z) * (math.pi / Lz)
        s, c = np.sin(arg), np.cos(arg)
        g = env[None, :] * s
        dg = denv[None, :] * s + env[None, :] * c * (np.outer(ll, np.ones_like(z)) * math.pi / Lz)
        return g, dg

    if device is not None:
        return _solenoidal_torch(nx, ny, nz, Lx, Lz, z_f, z_c, urms, U_bulk, chunk, device, mm,
                                 nn, kx, ky, A, G, dx, dy)

    def evaluate(coef_fn, zs, xshift, yshift):
        # value at x_i = i dx + xshift (i = 0..nx-1), y_j = j dy + yshift, z in zs
        out = np.empty((nx, ny, len(zs)))
        phase = np.exp(1j * (kx * xshift + ky * yshift))
        for z0 in range(0, len(zs), chunk):
            zz = zs[z0 : z0 + chunk]
            g, dg = G(zz)
            coef = coef_fn(g, dg) * phase[:, None]  # (M, nzc)
            C = np.zeros((nx, ny, len(zz)), dtype=np.complex128)
            np.add.at(C, (mm % nx, nn % ny), coef)
            out[:, :, z0 : z0 + len(zz)] = np.fft.ifft2(C, axes=(0, 1)).real * (nx * ny)
        return out

    ikx, iky = 1j * kx, 1j * ky
    # u = dAz/dy - dAy/dz ; v = dAx/dz - dAz/dx ; w = dAy/dx - dAx/dy
    fu = lambda g, dg: (iky * A["z"])[:, None] * g - A["y"][:, None] * dg  # noqa: E731
    fv = lambda g, dg: A["x"][:, None] * dg - (ikx * A["z"])[:, None] * g  # noqa: E731
    fw = lambda g, dg: (ikx * A["y"] - iky * A["x"])[:, None] * g  # noqa: E731

    zc_in = z_c[1 : nz + 1]
    zf_in = z_f[1:nz]
    # array index (i, j) <-> ghost index (i+1, j+1):
    #   u at x = I dx,         y = (J - 1/2) dy
    #   v at x = (I - 1/2) dx, y = J dy
    #   w at x = (I - 1/2) dx, y = (J - 1/2) dy
    up = evaluate(fu, zc_in, dx, 0.5 * dy)
    vp = evaluate(fv, zc_in, 0.5 * dx, dy)
    wp = evaluate(fw, zf_in, 0.5 * dx, 0.5 * dy)
    ss =(np.sum(up**2) + np.sum(vp**2) + np.sum(wp**2)) / (up.size + vp.size + wp.size)
    scale = urms / math.sqrt(ss) if ss > 0 else 0.0

    u = np.zeros((nx + 1, ny + 2, nz + 2))
    v = np.zeros((nx + 2, ny + 1, nz + 2))
    w = np.zeros((nx + 2, ny + 2, nz + 1))
    zeta = 2.0 * zc_in / Lz - 1.0
    base = 1.5 * U_bulk * (1.0 - zeta**2)
    u[1 : nx + 1, 1 : ny + 1, 1 : nz + 1] = scale * up + base[None, None, :]
    v[1 : nx + 1, 1 : ny + 1, 1 : nz + 1] = scale * vp
    w[1 : nx + 1, 1 : ny + 1, 1:nz] = scale * wp
    return u, v, w


def _solenoidal_torch(nx, ny, nz, Lx, Lz, z_f, z_c, urms, U_bulk, chunk, device, mm, nn, kx, ky, A,
                      G, dx, dy):
    import torch

    dev = torch.device(device)

    def upload(a):
        # host arrays go through pinned memory with an asynchronous copy, so
        # building an initial state never waits for work queued on the stream
        t = torch.from_numpy(np.ascontiguousarray(a))
        return t.pin_memory().to(dev, non_blocking=True)

    mi = upload(mm % nx)
    ni = upload(nn % ny)

    def evaluate(coef_fn, zs, xshift, yshift):
        out = torch.empty((nx, ny, len(zs)), dtype=torch.float64, device=dev)
        phase = np.exp(1j * (kx * xshift + ky * yshift))
        for z0 in range(0, len(zs), chunk):
            zz = zs[z0 : z0 + chunk]
            g, dg = G(zz)
            coef = upload(coef_fn(g, dg) * phase[:, None])
            C = torch.zeros((nx, ny, len(zz)), dtype=torch.complex128, device=dev)
            C.index_put_((mi, ni), coef, accumulate=True)
            out[:, :, z0 : z0 + len(zz)] = torch.fft.ifft2(C, dim=(0, 1)).real * (nx * ny)
        return out

    ikx, iky = 1j * kx, 1j * ky
    fu = lambda g, dg: (iky * A["z"])[:, None] * g - A["y"][:, None] * dg  # noqa: E731
    fv = lambda g, dg: A["x"][:, None] * dg - (ikx * A["z"])[:, None] * g  # noqa: E731
    fw = lambda g, dg: (ikx * A["y"] - iky * A["x"])[:, None] * g  # noqa: E731
    zc_in = z_c[1 : nz + 1]
    zf_in = z_f[1:nz]
    up = evaluate(fu, zc_in, dx, 0.5 * dy)
    vp = evaluate(fv, zc_in, 0.5 * dx, dy)
    wp = evaluate(fw, zf_in, 0.5 * dx, 0.5 * dy)
    # the scale stays on the device (no host synchronisation): the same IEEE
    # operations as urms / math.sqrt(sum / count) on the host
    ss = ((up**2).sum() + (vp**2).sum() + (wp**2).sum()) / (up.numel() + vp.numel()
                                                            + wp.numel())
    scale = torch.where(ss > 0, urms / torch.sqrt(ss), torch.zeros_like(ss))
    f64 = dict(dtype=torch.float64, device=dev)
    u = torch.zeros((nx + 1, ny + 2, nz + 2), **f64)
    v = torch.zeros((nx + 2, ny + 1, nz + 2), **f64)
    w = torch.zeros((nx + 2, ny + 2, nz + 1), **f64)
    zeta = 2.0 * zc_in / Lz - 1.0
    base = upload(1.5 * U_bulk * (1.0 - zeta**2))
    u[1 : nx + 1, 1 : ny + 1, 1 : nz + 1] = scale * up + base[None, None, :]
    v[1 : nx + 1, 1 : ny + 1, 1 : nz + 1] = scale * vp
    w[1 : nx + 1, 1 : ny + 1, 1:nz] = scale * wp
    return u, v, w
