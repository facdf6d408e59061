"""Analytic known-answer tests, run on the CPU oracle and on the HIP layer.

Ported from the reference's own test strategy (thresholds as there):
  zmap inverse            tests/test_zmap_inverse.py:22-46      < 1e-8
  interp convergence      tests/test_interp_convergence.py:50-59 ratio > 10 / > 40
  departure exactness     tests/test_departure_points.py:33-99  < 1e-13; ratio > 6
  shear advection         tests/test_sl_shear.py:50-66          < 1e-5; > 10; < cubic/10
  solid-body translation  tests/test_sl_solid_body.py:64-84     < 1e-12; > 12
  wall clamp              tests/test_sl_wall_clamp.py:20-45     clamped > 0, bound < 1.5
plus closed-form checks of the solvers: PCR and Thomas against dense LU, the
FFT-Poisson solve against a manufactured discrete solution.

The oracle rows pin the oracle by mathematics (independently of the golden
vectors); the hip rows (marked gpu) hold the HIP layer to the same bars.
"""

import math

import numpy as np
import pytest

from helpers import arrival, backend, make_field
from oracle import oracle as orc

BACKENDS = ["oracle", pytest.param("hip", marks=pytest.mark.gpu)]
TWO_PI = 2 * math.pi


def grid(gamma, nz, Lz=2.0, stretching="symmetric"):
    z_f, z_c, _, _ = orc.generate_grid(gamma, nz, Lz, stretching)
    return z_f, z_c


def test_zmap_inverse(check):
    """Faces invert to integers; random points round-trip through the map."""
    Lz = 2.0
    for stretching in ("symmetric", "bottom"):
        for gamma in (0.8, 1.8, 2.6):
            for nz in (64, 180):
                z_f, z_c = grid(gamma, nz, Lz, stretching)
                sl = orc.SLOracle(4, 4, nz, 0.5, 0.5, 2.0, 2.0, Lz, z_f, z_c, gamma,
                                  stretching_type=stretching)
                kf = sl.face_coord(z_f)
                err_f = np.abs(kf - np.arange(nz + 1)).max()
                rng = np.random.default_rng(0)
                kr = rng.random(5000) * nz
                g = np.float64(gamma)
                if stretching == "symmetric":
                    zr = 0.5 * Lz * (1 + np.tanh(g * (2.0 * kr / nz - 1.0)) / np.tanh(g))
                else:
                    zr = Lz * (1.0 - np.tanh(g * (1.0 - kr / nz)) / np.tanh(g))
                err_r = np.abs(sl.face_coord(zr) - kr).max()
                check(f"zmap {stretching} gamma={gamma} nz={nz}", err_f < 1e-8 and err_r < 1e-8,
                      f"face={err_f:.1e} rand={err_r:.1e}")


def test_interp_convergence(check):
    Lx = Ly = TWO_PI
    Lz, gamma = 2.0, 1.8

    def f(x, y, z):
        return (np.sin(TWO_PI * x / Lx + 0.3) * np.cos(2 * TWO_PI * y / Ly - 1.1)
                * np.sin(1.5 * math.pi * z / Lz + 0.4))

    def err(n, order):
        z_f, z_c = grid(gamma, n, Lz)
        sl = orc.SLOracle(n, n, n, Lx / n, Ly / n, Lx, Ly, Lz, z_f, z_c, gamma, order=order)
        u = make_field("u", f, n, n, n, Lx / n, Ly / n, z_f, z_c)
        rng = np.random.default_rng(1234)
        xq, yq = rng.random(20000) * Lx, rng.random(20000) * Ly
        zq = sl.z_lo + rng.random(20000) * (sl.z_hi - sl.z_lo)
        return np.abs(sl.interp("u", u, xq, yq, zq) - f(xq, yq, zq)).max()

    for order, min_ratio in ((4, 10.0), (6, 40.0)):
        e = [err(n, order) for n in (32, 64, 128)]
        r1, r2 = e[0] / e[1], e[1] / e[2]
        check(f"interp order={order}", r1 > min_ratio and r2 > min_ratio,
              f"errors={[f'{x:.2e}' for x in e]} ratios={r1:.1f},{r2:.1f}")


def _mids(nx, ny, nz, dx, dy, z_f, z_c, fu, fv, fw):
    return (make_field("u", fu, nx, ny, nz, dx, dy, z_f, z_c),
            make_field("v", fv, nx, ny, nz, dx, dy, z_f, z_c),
            make_field("w", fw, nx, ny, nz, dx, dy, z_f, z_c))


@pytest.mark.parametrize("be", BACKENDS)
def test_departure_points(check, be):
    Lx = Ly = TWO_PI
    Lz, gamma, nx, ny, nz = 2.0, 1.5, 16, 16, 32
    dx, dy = Lx / nx, Ly / ny
    z_f, z_c = grid(gamma, nz, Lz)
    b = backend(be, nx, ny, nz, Lx, Ly, Lz, z_f, z_c, gamma)
    U0, V0, W0 = 0.7, -0.3, 0.02
    mids = _mids(nx, ny, nz, dx, dy, z_f, z_c, lambda X, Y, Z: U0 + 0 * X,
                 lambda X, Y, Z: V0 + 0 * X, lambda X, Y, Z: W0 + 0 * X)
    dt = 0.05
    for comp in "uvw":
        xa, ya, za = arrival(comp, nx, ny, nz, dx, dy, z_f, z_c)
        xd, yd, zd = b.departure(comp, *mids, dt)
        e = max(np.abs(xd - (xa - dt * U0)).max(), np.abs(yd - (ya - dt * V0)).max(),
                np.abs(zd - (za - dt * W0)).max())
        check(f"uniform translation {comp}", e < 1e-13, f"err={e:.1e}")
    S = 0.8
    mids = _mids(nx, ny, nz, dx, dy, z_f, z_c, lambda X, Y, Z: S * Z, lambda X, Y, Z: 0 * X,
                 lambda X, Y, Z: 0 * X)
    xa, ya, za = arrival("u", nx, ny, nz, dx, dy, z_f, z_c)
    xd, yd, zd = b.departure("u", *mids, dt)
    e = max(np.abs(xd - (xa - dt * S * za)).max(), np.abs(yd - ya).max(), np.abs(zd - za).max())
    check("linear shear", e < 1e-13, f"err={e:.1e}")
    x0, z0, om = math.pi, 1.0, 0.3
    errs = []
    for dtv in (0.2, 0.1, 0.05):
        b8 = backend(be, nx, ny, nz, Lx, Ly, Lz, z_f, z_c, gamma, n_traj_iters=8)
        mids = _mids(nx, ny, nz, dx, dy, z_f, z_c, lambda X, Y, Z: -om * (Z - z0) + 0 * X,
                     lambda X, Y, Z: 0 * X, lambda X, Y, Z: om * (X - x0) + 0 * Z)
        xa, ya, za = arrival("u", nx, ny, nz, dx, dy, z_f, z_c)
        xd, yd, zd = b8.departure("u", *mids, dtv)
        th = om * dtv
        xe = x0 + math.cos(th) * (xa - x0) + math.sin(th) * (za - z0)
        ze = z0 - math.sin(th) * (xa - x0) + math.cos(th) * (za - z0)
        mask = np.broadcast_to((np.abs(za - z0) < 0.7) & (xa > 1.0) & (xa < Lx - 1.0),
                               (nx, ny, nz))
        errs.append(np.maximum(np.abs(xd - xe), np.abs(zd - ze))[mask].max())
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    check("rotation local error O(dt^3)", r1 > 6.0 and r2 > 6.0,
          f"errors={[f'{x:.2e}' for x in errs]} ratios={r1:.1f},{r2:.1f}")


def _shear_error(be, nx, order):
    Lx = Ly = TWO_PI
    Lz, gamma, S, DT = 2.0, 1.5, 0.8, 0.1
    ny, nz = nx, 32
    dx, dy = Lx / nx, Ly / ny
    z_f, z_c = grid(gamma, nz, Lz)
    b = backend(be, nx, ny, nz, Lx, Ly, Lz, z_f, z_c, gamma, order=order)

    def f(X, Y, Z):
        return np.sin(TWO_PI * X / Lx) + 0 * Y + 0 * Z

    flds = [make_field(c, f, nx, ny, nz, dx, dy, z_f, z_c) for c in "uvw"]
    mu = make_field("u", lambda X, Y, Z: S * Z + 0 * X, nx, ny, nz, dx, dy, z_f, z_c)
    mv = np.zeros((nx + 2, ny + 1, nz + 2))
    mw = np.zeros((nx + 2, ny + 2, nz + 1))
    out, _ = b.advect(*flds, mu, mv, mw, DT)
    errs = []
    for comp, o in zip("uvw", out):
        xa, ya, za = arrival(comp, nx, ny, nz, dx, dy, z_f, z_c)
        ex = np.sin(TWO_PI * (xa - S * za * DT) / Lx) + 0 * ya
        got = o[1 : nx + 1, 1 : ny + 1, 1:nz] if comp == "w" else o[1 : nx + 1, 1 : ny + 1,
                                                                      1 : nz + 1]
        errs.append(np.abs(got - ex).max())
    return max(errs)


@pytest.mark.parametrize("be", BACKENDS)
def test_sl_shear(check, be):
    e_c = _shear_error(be, 64, 4)
    check("shear, tricubic nx=64", e_c < 1e-5, f"err={e_c:.3e}")
    e1, e2 = _shear_error(be, 32, 4), e_c
    check("shear convergence O(h^4)", e1 / e2 > 10.0, f"{e1:.3e},{e2:.3e} ratio={e1 / e2:.1f}")
    e_q = _shear_error(be, 64, 6)
    check("triquintic beats tricubic", e_q < e_c / 10, f"quintic={e_q:.3e} cubic={e_c:.3e}")


def _solid_body(be, n, fn, dt, order=4):
    Lx = Ly = TWO_PI
    Lz, gamma, U0, V0 = 2.0, 1.5, 0.9, -0.4
    dx, dy = Lx / n, Ly / n
    z_f, z_c = grid(gamma, n, Lz)
    b = backend(be, n, n, n, Lx, Ly, Lz, z_f, z_c, gamma, order=order)
    flds = [make_field(c, fn, n, n, n, dx, dy, z_f, z_c) for c in "uvw"]
    mids = (make_field("u", lambda X, Y, Z: U0 + 0 * X, n, n, n, dx, dy, z_f, z_c),
            make_field("v", lambda X, Y, Z: V0 + 0 * X, n, n, n, dx, dy, z_f, z_c),
            np.zeros((n + 2, n + 2, n + 1)))
    out, ncl = b.advect(*flds, *mids, dt)
    errs = []
    for comp, o in zip("uvw", out):
        xa, ya, za = arrival(comp, n, n, n, dx, dy, z_f, z_c)
        ex = fn(xa - dt * U0, ya - dt * V0, za) + 0 * (xa + ya + za)
        got = o[1 : n + 1, 1 : n + 1, 1:n] if comp == "w" else o[1 : n + 1, 1 : n + 1, 1 : n + 1]
        errs.append(np.abs(got - ex).max())
    return max(errs), ncl


@pytest.mark.parametrize("be", BACKENDS)
def test_sl_solid_body(check, be):
    e, ncl = _solid_body(be, 16, lambda X, Y, Z: 1.0 + 0.5 * Z - 0.3 * Z**2 + 0.1 * Z**3 + 0 * X,
                         0.08)
    check("cubic-in-z exactness", e < 1e-12 and ncl == 0, f"err={e:.2e} clamped={ncl}")

    def f(X, Y, Z):
        return (np.sin(TWO_PI * X / TWO_PI) * np.cos(TWO_PI * Y / TWO_PI)
                * np.sin(1.2 * math.pi * Z / 2.0 + 0.2))

    e1, _ = _solid_body(be, 24, f, 0.08)
    e2, _ = _solid_body(be, 48, f, 0.04)
    check("smooth-field convergence O(h^4)", e1 / e2 > 12.0,
          f"{e1:.3e},{e2:.3e} ratio={e1 / e2:.1f}")


@pytest.mark.parametrize("be", BACKENDS)
def test_sl_wall_clamp(check, be):
    Lx = Ly = TWO_PI
    Lz, gamma, n, nz = 2.0, 2.0, 16, 48
    dx, dy = Lx / n, Ly / n
    z_f, z_c = grid(gamma, nz, Lz)

    def f(X, Y, Z):
        return np.cos(TWO_PI * X / Lx) * np.sin(math.pi * Z / Lz) + 0 * Y

    for w0, label in ((-0.5, "toward top wall"), (0.5, "toward bottom wall")):
        b = backend(be, n, n, nz, Lx, Ly, Lz, z_f, z_c, gamma)
        flds = [make_field(c, f, n, n, nz, dx, dy, z_f, z_c) for c in "uvw"]
        mids = (np.zeros((n + 1, n + 2, nz + 2)), np.zeros((n + 2, n + 1, nz + 2)),
                np.full((n + 2, n + 2, nz + 1), w0))
        out, ncl = b.advect(*flds, *mids, 0.5)
        finite = all(np.isfinite(o).all() for o in out)
        bound = max(np.abs(o).max() for o in out)
        check(f"wall clamp {label}", finite and ncl > 0 and bound < 1.5,
              f"clamped={ncl} max|field|={bound:.3f}")


def test_tridiagonal_solvers_vs_dense(check):
    rng = np.random.default_rng(5)
    n, batch = 23, 7
    a = rng.normal(size=n)
    c = rng.normal(size=n)
    b = 4.0 + np.abs(a) + np.abs(c) + rng.random(n)
    d = rng.normal(size=(batch, n))
    A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    ref = np.linalg.solve(A, d.T).T
    x = orc.pcr_solve(a, b, c, d)
    check("PCR vs dense", np.abs(x - ref).max() < 1e-13, f"{np.abs(x - ref).max():.1e}")
    A2 = np.broadcast_to(a, (batch, n)).copy()
    x2 = orc.solve_tridiagonal(A2, np.broadcast_to(b, (batch, n)).copy(),
                               np.broadcast_to(c, (batch, n)).copy(), d + 1j * d[::-1])
    ref2 = np.linalg.solve(A, (d + 1j * d[::-1]).T).T
    check("Thomas (complex rhs) vs dense", np.abs(x2 - ref2).max() < 1e-13,
          f"{np.abs(x2 - ref2).max():.1e}")


def _discrete_laplacian(p, nx, ny, nz, dx, dy, dz_c, dz_f):
    """Staggered Laplacian of a ghost-filled p at the cell centres."""
    c = p[1 : nx + 1, 1 : ny + 1, 1 : nz + 1]
    lx = (p[2 : nx + 2, 1 : ny + 1, 1 : nz + 1] - 2 * c + p[0:nx, 1 : ny + 1, 1 : nz + 1]) / dx**2
    ly = (p[1 : nx + 1, 2 : ny + 2, 1 : nz + 1] - 2 * c + p[1 : nx + 1, 0:ny, 1 : nz + 1]) / dy**2
    up = (p[1 : nx + 1, 1 : ny + 1, 2 : nz + 2] - c) / dz_c[1 : nz + 1]
    dn = (c - p[1 : nx + 1, 1 : ny + 1, 0:nz]) / dz_c[0:nz]
    return lx + ly + (up - dn) / dz_f


def test_poisson_manufactured(check):
    """Apply the discrete Laplacian (the one projection.py:24-114 inverts:
    modified wavenumbers = central differences, Neumann ghosts) to a random
    p, solve, and recover p up to the constant of the (0,0) mode."""
    nx, ny, nz = 12, 10, 16
    Lx, Ly = 4 * math.pi, 2 * math.pi
    dx, dy = Lx / nx, Ly / ny
    z_f, z_c, dz_f, dz_c = orc.generate_grid(1.5, nz, 2.0)
    rng = np.random.default_rng(2)
    p = np.zeros((nx + 2, ny + 2, nz + 2))
    p[1 : nx + 1, 1 : ny + 1, 1 : nz + 1] = rng.normal(size=(nx, ny, nz))
    p[0], p[nx + 1] = p[nx], p[1]
    p[:, 0], p[:, ny + 1] = p[:, ny], p[:, 1]
    p[:, :, 0], p[:, :, nz + 1] = p[:, :, 1], p[:, :, nz]
    rhs = _discrete_laplacian(p, nx, ny, nz, dx, dy, dz_c, dz_f)
    fft = orc.initialize_fft_solver(nx, ny, nz, dx, dy, dz_c, dz_f)
    q = orc.solve_poisson_fft(rhs, fft)
    diff = (q - p)[1 : nx + 1, 1 : ny + 1, 1 : nz + 1]
    e = np.abs(diff - diff.mean()).max() / np.abs(p).max()
    check("recovered p (mod a constant)", e < 1e-11, f"rel={e:.1e}")
    r2 = _discrete_laplacian(q, nx, ny, nz, dx, dy, dz_c, dz_f)
    e2 = np.abs(r2 - rhs).max() / np.abs(rhs).max()
    check("residual of the solve", e2 < 1e-11, f"rel={e2:.1e}")
