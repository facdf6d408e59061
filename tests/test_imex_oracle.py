"""Pins the oracle's Eulerian IMEX right-hand side
(oracle.compute_momentum_rhs_fused_imex, operators.py:639-835).

* Uniform flow: every flux difference and the Laplacian vanish exactly.
* Flux form: with periodic x/y ghosts and w = 0 on the walls, the
  volume-weighted sum of the u- and v-advection telescopes to rounding in
  all three directions (dropping the periodic face nx / ny, the leak the
  reference's comments warn about, fails this).
* xy-diffusion of a single y-mode: the discrete Laplacian eigenvalue
  nu (2 cos(theta) - 2) / dy^2, exactly.
* Second-order convergence of all three components to the analytic RHS of
  a smooth manufactured field on the stretched grid (ghosts filled with the
  analytic field): error ratio ~4 per halving.
"""

import math

import numpy as np

import oracle.oracle as orc


def _grid(n, nz=None, gamma=1.5, L=2 * math.pi, Lz=2.0):
    nz = nz or n
    z_f, z_c, dz_f, dz_c = orc.generate_grid(gamma, nz, Lz)
    return dict(nx=n, ny=n, nz=nz, dx=L / n, dy=L / n, Lz=Lz, z_f=np.asarray(z_f),
                z_c=np.asarray(z_c), dz_f=np.asarray(dz_f), dz_c=np.asarray(dz_c))


def _rhs(g, u, v, w, nu):
    return orc.compute_momentum_rhs_fused_imex(u, v, w, g["nx"], g["ny"], g["nz"], g["dx"],
                                               g["dy"], g["dz_c"], g["dz_f"], nu)


def _shapes(g):
    nx, ny, nz = g["nx"], g["ny"], g["nz"]
    return (nx + 1, ny + 2, nz + 2), (nx + 2, ny + 1, nz + 2), (nx + 2, ny + 2, nz + 1)


def test_uniform_flow_rhs_zero(check):
    g = _grid(8, 10)
    su, sv, sw = _shapes(g)
    ru, rv, rw = _rhs(g, np.full(su, 0.7), np.full(sv, -0.3), np.zeros(sw), 1e-3)
    check("uniform flow: rhs == 0 exactly",
          not ru.any() and not rv.any() and not rw.any(),
          f"{np.abs(ru).max()} {np.abs(rv).max()} {np.abs(rw).max()}")


def test_flux_form_telescopes(check):
    g = _grid(12, 10)
    nx, ny, nz = g["nx"], g["ny"], g["nz"]
    rng = np.random.default_rng(3)
    u, v, w = (rng.standard_normal(s) for s in _shapes(g))
    orc.apply_bc_all(u, v, w)
    ru, rv, _ = _rhs(g, u, v, w, 0.0)
    vol = g["dx"] * g["dy"] * g["dz_f"][None, None, :]
    scale = np.abs(ru[1:nx + 1, 1:ny + 1, 1:nz + 1]).max()
    su = float(np.sum(ru[1:nx + 1, 1:ny + 1, 1:nz + 1] * vol))
    sv = float(np.sum(rv[1:nx + 1, 1:ny + 1, 1:nz + 1] * vol))
    tol = 1e-13 * scale * vol.sum()
    check("u momentum: advection sums to zero", abs(su) < tol, f"{su:.2e} (tol {tol:.1e})")
    check("v momentum: advection sums to zero", abs(sv) < tol, f"{sv:.2e}")
    # a flux form that skipped the periodic face nx would leak: show the test sees it
    ru_cut = ru.copy()
    ru_cut[nx] = 0.0
    leak = float(np.sum(ru_cut[1:nx + 1, 1:ny + 1, 1:nz + 1] * vol))
    check("dropping face nx is visible (sanity)", abs(leak) > 1e3 * tol, f"{leak:.2e}")


def test_xy_diffusion_of_a_mode(check):
    g = _grid(12, 6)
    nx, ny, nz = g["nx"], g["ny"], g["nz"]
    m = 3
    th = 2 * math.pi * m / ny
    f = np.cos(th * np.arange(ny + 2))
    su, sv, sw = _shapes(g)
    u = np.broadcast_to(f[None, :, None], su).copy()
    nu = 0.02
    ru, rv, rw = _rhs(g, u, np.zeros(sv), np.zeros(sw), nu)
    exp = nu * (2 * math.cos(th) - 2) / g["dy"] ** 2 * u[1:nx + 1, 1:ny + 1, 1:nz + 1]
    err = np.abs(ru[1:nx + 1, 1:ny + 1, 1:nz + 1] - exp).max()
    check("u-rhs = nu (2 cos th - 2)/dy^2 u", err < 1e-14 * np.abs(exp).max(), f"{err:.1e}")
    check("v, w rhs zero", not rv.any() and not rw.any(), "")


def _manufactured(g, nu):
    """Fields evaluated at every entry's position (ghosts included) and the
    analytic rhs = nu (f_xx + f_yy) - div(f V) at the interior nodes."""
    nx, ny, nz, dx, dy, Lz = g["nx"], g["ny"], g["nz"], g["dx"], g["dy"], g["Lz"]
    k = math.pi / Lz

    def AB(x, y):
        A = np.sin(x) * (1 + 0.1 * np.cos(y))
        Ax, Ay = np.cos(x) * (1 + 0.1 * np.cos(y)), -0.1 * np.sin(x) * np.sin(y)
        Axx, Ayy = -A, -0.1 * np.sin(x) * np.cos(y)
        B = 0.3 * np.cos(x) * np.sin(y)
        Bx, By = -0.3 * np.sin(x) * np.sin(y), 0.3 * np.cos(x) * np.cos(y)
        Bxx, Byy = -B, -B
        C = 0.2 * np.sin(x + y)
        Cx = Cy = 0.2 * np.cos(x + y)
        Cxx = Cyy = -C
        return A, Ax, Ay, Axx, Ayy, B, Bx, By, Bxx, Byy, C, Cx, Cy, Cxx, Cyy

    def hz(z):
        return 1 + 0.5 * np.cos(k * z), -0.5 * k * np.sin(k * z)

    def sz(z):
        return np.sin(k * z), k * np.cos(k * z)

    def at(x, y, z):
        A, Ax, Ay, Axx, Ayy, B, Bx, By, Bxx, Byy, C, Cx, Cy, Cxx, Cyy = AB(x, y)
        h, hp = hz(z)
        s, sp = sz(z)
        u, v, w = A * h, B * h, C * s
        ux, uy, uz = Ax * h, Ay * h, A * hp
        vx, vy, vz = Bx * h, By * h, B * hp
        wx, wy, wz = Cx * s, Cy * s, C * sp
        ru = nu * (Axx + Ayy) * h - (2 * u * ux + (vy * u + v * uy) + (wz * u + w * uz))
        rv = nu * (Bxx + Byy) * h - ((ux * v + u * vx) + 2 * v * vy + (wz * v + w * vz))
        rw = nu * (Cxx + Cyy) * s - ((ux * w + u * wx) + (vy * w + v * wy) + 2 * w * wz)
        return u, v, w, ru, rv, rw

    xi = np.arange(nx + 2)
    yj = np.arange(ny + 2)
    zc, zf = g["z_c"], g["z_f"]
    X, Y, Z = np.meshgrid(xi[:nx + 1] * dx, (yj - 0.5) * dy, zc, indexing="ij")
    u, _, _, ru, _, _ = at(X, Y, Z)
    X, Y, Z = np.meshgrid((xi - 0.5) * dx, yj[:ny + 1] * dy, zc, indexing="ij")
    _, v, _, _, rv, _ = at(X, Y, Z)
    X, Y, Z = np.meshgrid((xi - 0.5) * dx, (yj - 0.5) * dy, zf, indexing="ij")
    _, _, w, _, _, rw = at(X, Y, Z)
    return u, v, w, ru, rv, rw


def test_second_order_convergence(check):
    nu = 0.01
    errs = []
    for n in (16, 32, 64):
        g = _grid(n, n, gamma=1.2)
        nx, ny, nz = n, n, n
        u, v, w, eu, ev, ew = _manufactured(g, nu)
        ru, rv, rw = _rhs(g, u, v, w, nu)
        I, J, K, Kw = slice(1, nx + 1), slice(1, ny + 1), slice(1, nz + 1), slice(1, nz)
        errs.append((np.abs(ru[I, J, K] - eu[I, J, K]).max(),
                     np.abs(rv[I, J, K] - ev[I, J, K]).max(),
                     np.abs(rw[I, J, Kw] - ew[I, J, Kw]).max()))
    for c, name in enumerate("uvw"):
        r1, r2 = errs[0][c] / errs[1][c], errs[1][c] / errs[2][c]
        check(f"{name}: second order (error ratios ~4)", 3.3 < r2 < 4.7 and 3.0 < r1 < 5.0,
              f"errors {[f'{e[c]:.2e}' for e in errs]} ratios {r1:.2f} {r2:.2f}")
