"""Seeded solenoidal initial fields for the channel: numpy on the host, or on
the device through slc_init_solenoidal (the plan's library).

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
    """Integer triples (m, n, l) with kmin <= |(m, n, l)| <= kmax, |m|, |n| <= K
    (K = floor(kmax), capped below the Nyquist mode), 1 <= l <= floor(kmax);
    m outermost, l innermost.  Returns (m, n, l, |k|) arrays."""
    K = int(math.floor(kmax))
    K = min(K, nx // 2 - 1, ny // 2 - 1)
    L = int(math.floor(kmax))
    m, n, l = np.meshgrid(np.arange(-K, K + 1), np.arange(-K, K + 1), np.arange(1, L + 1),
                          indexing="ij")
    m, n, l = m.ravel(), n.ravel(), l.ravel()
    k = np.sqrt((m * m + n * n + l * l).astype(np.float64))
    keep = (kmin <= k) & (k <= kmax)
    return m[keep], n[keep], l[keep], k[keep], K, L


def solenoidal_fields(nx, ny, nz, Lx, Ly, Lz, z_f, z_c, seed, urms=0.1, kmin=1.0, kmax=8.0,
                      spectrum="flat", U_bulk=1.0, chunk=32, plan=None):
    """Return ghost-shaped arrays (u, v, w) with interiors set: numpy arrays
    (x-y synthesis by numpy FFTs), or, when a plan is given, float64 device
    tensors on the plan's device built by slc_init_solenoidal (direct
    summation over the same modes and amplitudes; rounding aside the same
    fields)."""
    z_f = np.asarray(z_f, dtype=np.float64)
    z_c = np.asarray(z_c, dtype=np.float64)
    rng = np.random.Generator(np.random.PCG64(seed))
    mm, nn, ll, kk, K, L = _modes(kmin, kmax, nx, ny)
    if mm.size == 0:
        raise ValueError("no Fourier modes in the requested band")
    p = {"flat": 1.0, "k-5/3": 17.0 / 6.0}[spectrum]
    M = mm.size
    amp = kk ** (-p)
    A = {c: (rng.standard_normal(M) + 1j * rng.standard_normal(M)) * amp for c in "xyz"}
    dx, dy = Lx / nx, Ly / ny

    def G(z, ls=ll):
        # G_l(z) and dG_l/dz for each l of ls (rows) at each z (columns)
        zeta = 2.0 * z / Lz - 1.0
        env = (1.0 - zeta**2) ** 2
        denv = -8.0 * zeta * (1.0 - zeta**2) / Lz
        arg = np.outer(ls, z) * (math.pi / Lz)
        s, c = np.sin(arg), np.cos(arg)
        g = env[None, :] * s
        dg = denv[None, :] * s + env[None, :] * c * (np.outer(ls, np.ones_like(z)) * math.pi / Lz)
        return g, dg

    if plan is not None:
        return _solenoidal_hip(plan, nx, ny, nz, Lx, Ly, Lz, z_f, z_c, urms, U_bulk, mm, nn, ll,
                               K, L, A, G, dx, dy)
    kx = 2 * math.pi * mm / Lx
    ky = 2 * math.pi * nn / Ly

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


_HIP_TABLES = {}


def _solenoidal_hip(plan, nx, ny, nz, Lx, Ly, Lz, z_f, z_c, urms, U_bulk, mm, nn, ll, K, L, A,
                    G, dx, dy):
    """The fields on the plan's device with slc_init_solenoidal.  Per request
    the host draws nothing new here (the amplitudes A are given) and uploads
    the per-bin coefficients of G_l and dG_l/dz (five (W^2, L) complex blocks,
    one pinned copy); the z tables and the staggered-position phase factors
    depend only on the grid and the mode box and are built once per grid.
    Four kernel launches, no host synchronisation."""
    import torch

    import hashlib

    dev = plan.device
    W = 2 * K + 1
    zdig = hashlib.sha1(np.ascontiguousarray(z_c, dtype=np.float64).tobytes()
                        + np.ascontiguousarray(z_f, dtype=np.float64).tobytes()).hexdigest()
    key = (str(dev), nx, ny, nz, Lx, Ly, Lz, zdig, K, L, float(U_bulk))
    tabs = _HIP_TABLES.get(key)
    if tabs is None:
        zc_in = np.asarray(z_c[1 : nz + 1], dtype=np.float64)
        zf_in = np.asarray(z_f[1:nz], dtype=np.float64)
        ls = np.arange(1, L + 1)
        gc, dgc = G(zc_in, ls)
        gf, _ = G(zf_in, ls)
        zeta = 2.0 * zc_in / Lz - 1.0
        base = 1.5 * U_bulk * (1.0 - zeta**2)
        zt = np.concatenate([gc.ravel(), dgc.ravel(), gf.ravel(), base])
        ks = np.arange(-K, K + 1)
        kx, ky = 2 * math.pi * ks / Lx, 2 * math.pi * ks / Ly
        xi, yj = np.arange(nx) * dx, np.arange(ny) * dy
        ph = []
        # u at (i dx + dx, j dy + dy/2), v at (i dx + dx/2, j dy + dy), w at
        # (i dx + dx/2, j dy + dy/2): array index i <-> ghost index i + 1
        for sx, sy in ((dx, 0.5 * dy), (0.5 * dx, dy), (0.5 * dx, 0.5 * dy)):
            ph.append(np.exp(1j * np.outer(kx, xi + sx)).ravel().view(np.float64))
            ph.append(np.exp(1j * np.outer(ky, yj + sy)).ravel().view(np.float64))
        tabs = (torch.from_numpy(zt).to(dev), torch.from_numpy(np.concatenate(ph)).to(dev))
        _HIP_TABLES[key] = tabs
    ztd, phd = tabs

    def dense(a):
        d = np.zeros((W, W, L), dtype=np.complex128)
        d[mm + K, nn + K, ll - 1] = a
        return d.reshape(W * W, L)

    ks = np.arange(-K, K + 1)
    ikx = (1j * np.repeat(2 * math.pi * ks / Lx, W))[:, None]   # bin b = (m + K) W + (n + K)
    iky = (1j * np.tile(2 * math.pi * ks / Ly, W))[:, None]
    Ax, Ay, Az = dense(A["x"]), dense(A["y"]), dense(A["z"])
    # u = dAz/dy - dAy/dz ; v = dAx/dz - dAz/dx ; w = dAy/dx - dAx/dy: the
    # coefficients of g_l (alpha) and of dg_l/dz (beta) per bin and l
    ab = np.concatenate([x.ravel() for x in (iky * Az, -Ay, -ikx * Az, Ax, ikx * Ay - iky * Ax)])
    buf = torch.from_numpy(ab.view(np.float64))
    abd = buf.pin_memory().to(dev, non_blocking=True) if dev.type == "cuda" else buf.to(dev)
    f64 = dict(dtype=torch.float64, device=dev)
    # scratch from the caching allocator, returned when the call's work is done
    work = torch.empty((plan.init_solenoidal_work(W) + 7) // 8, **f64)
    u = torch.zeros((nx + 1, ny + 2, nz + 2), **f64)
    v = torch.zeros((nx + 2, ny + 1, nz + 2), **f64)
    w = torch.zeros((nx + 2, ny + 2, nz + 1), **f64)
    plan.init_solenoidal(W, L, abd, ztd, phd, urms, work, u, v, w)
    return u, v, w
