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
                      spectrum="flat", U_bulk=1.0, chunk=32, device=None):
    """Return ghost-shaped arrays (u, v, w) with interiors set: numpy arrays,
    or float64 torch tensors on `device` when one is given (the x-y
    synthesis then runs as torch FFTs there; same modes and amplitudes, FFT
    rounding aside)."""
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

    if device is not None:
        return _solenoidal_torch(nx, ny, nz, Lx, Ly, Lz, z_f, z_c, urms, U_bulk, device, mm, nn,
                                 ll, K, L, A, G, dx, dy)
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


def _solenoidal_torch(nx, ny, nz, Lx, Ly, Lz, z_f, z_c, urms, U_bulk, device, mm, nn, ll, K, L,
                      A, G, dx, dy):
    """The same fields on the device: the amplitudes go into dense
    (m, n, l) arrays in one upload; per component and z chunk, one complex
    matrix product sums the l modes of every (m, n) bin, the bins are placed
    in the spectral plane (no two bins share a wavenumber, so nothing is
    accumulated and the sums do not depend on launch order), and one batched
    inverse FFT gives the planes.  No host synchronisation."""
    import torch

    dev = torch.device(device)
    W = 2 * K + 1
    ks = np.arange(-K, K + 1)
    kxb = np.repeat(2 * math.pi * ks / Lx, W)   # bin b = (m + K) W + (n + K)
    kyb = np.tile(2 * math.pi * ks / Ly, W)

    def dense(a):
        d = np.zeros((W, W, L), dtype=np.complex128)
        d[mm + K, nn + K, ll - 1] = a
        return d.reshape(W * W, L)

    Ax, Ay, Az = dense(A["x"]), dense(A["y"]), dense(A["z"])
    ikx, iky = (1j * kxb)[:, None], (1j * kyb)[:, None]
    # u = dAz/dy - dAy/dz ; v = dAx/dz - dAz/dx ; w = dAy/dx - dAx/dy: the
    # coefficients of g_l (alpha) and of dg_l/dz (beta) per bin and l
    alpha_beta = [iky * Az, -Ay, -ikx * Az, Ax, ikx * Ay - iky * Ax]
    zc_in = np.asarray(z_c[1 : nz + 1], dtype=np.float64)
    zf_in = np.asarray(z_f[1:nz], dtype=np.float64)
    gc, dgc = G(zc_in, np.arange(1, L + 1))
    gf, _ = G(zf_in, np.arange(1, L + 1))
    zeta = 2.0 * zc_in / Lz - 1.0
    base = 1.5 * U_bulk * (1.0 - zeta**2)
    shifts = [(dx, 0.5 * dy), (0.5 * dx, dy), (0.5 * dx, 0.5 * dy)]  # u, v, w positions
    phases = [np.exp(1j * (kxb * sx + kyb * sy)) for sx, sy in shifts]
    bins = np.concatenate([ks.repeat(W) % nx, np.tile(ks, W) % ny]).astype(np.float64)
    # one pinned upload of everything the synthesis reads
    cplx = alpha_beta + phases
    host = np.concatenate([x.ravel().view(np.float64) for x in cplx]
                          + [gc.ravel(), dgc.ravel(), gf.ravel(), base, bins])
    buf = torch.from_numpy(host)
    buf = buf.pin_memory().to(dev, non_blocking=True) if dev.type == "cuda" else buf.to(dev)
    parts, off = [], 0
    for x in cplx:
        parts.append(torch.view_as_complex(buf[off:off + 2 * x.size].view(-1, 2)).view(x.shape))
        off += 2 * x.size
    au, bu, av, bv, aw, pu, pv, pw = parts
    gct = buf[off:off + gc.size].view(gc.shape); off += gc.size
    dgct = buf[off:off + dgc.size].view(dgc.shape); off += dgc.size
    gft = buf[off:off + gf.size].view(gf.shape); off += gf.size
    baset = buf[off:off + nz]; off += nz
    mi = buf[off:off + W * W].long(); off += W * W
    ni = buf[off:off + W * W].long()
    # z chunks of at most 2^27 plane points (2 GiB of complex spectrum)
    chunk = max(1, min(nz, (1 << 27) // (nx * ny)))

    def evaluate(alpha, beta, g, dg, phase):
        nzs = g.shape[1]
        out = torch.empty((nx, ny, nzs), dtype=torch.float64, device=dev)
        for z0 in range(0, nzs, chunk):
            z1 = min(nzs, z0 + chunk)
            coef = alpha @ g[:, z0:z1].to(torch.complex128)
            if beta is not None:
                coef = coef + beta @ dg[:, z0:z1].to(torch.complex128)
            C = torch.zeros((nx, ny, z1 - z0), dtype=torch.complex128, device=dev)
            C[mi, ni] = coef * phase[:, None]
            out[:, :, z0:z1] = torch.fft.ifft2(C, dim=(0, 1)).real * (nx * ny)
        return out

    up = evaluate(au, bu, gct, dgct, pu)
    vp = evaluate(av, bv, gct, dgct, pv)
    wp = evaluate(aw, None, gft, None, pw)
    # the scale stays on the device (no host synchronisation): the same IEEE
    # operations as urms / math.sqrt(sum / count) on the host
    ss = ((up**2).sum() + (vp**2).sum() + (wp**2).sum()) / (up.numel() + vp.numel()
                                                            + wp.numel())
    scale = torch.where(ss > 0, urms / torch.sqrt(ss), torch.zeros_like(ss))
    f64 = dict(dtype=torch.float64, device=dev)
    u = torch.zeros((nx + 1, ny + 2, nz + 2), **f64)
    v = torch.zeros((nx + 2, ny + 1, nz + 2), **f64)
    w = torch.zeros((nx + 2, ny + 2, nz + 1), **f64)
    u[1 : nx + 1, 1 : ny + 1, 1 : nz + 1] = scale * up + baset[None, None, :]
    v[1 : nx + 1, 1 : ny + 1, 1 : nz + 1] = scale * vp
    w[1 : nx + 1, 1 : ny + 1, 1:nz] = scale * wp
    return u, v, w
