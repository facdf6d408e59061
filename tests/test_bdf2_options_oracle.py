"""The two non-default BDF2 options of the oracle (solver.py:256-281, 846-1012),
pinned by properties that do not restate their formulas:

  lagged xy RHS   with V^{n-1} = V^n the lagged and extrapolated predictors
                  are the same doubles (2R - R = R exactly); with V^{n-1} = 0
                  they differ by dt_eff R^n, R^n = nu (d2x + d2y) V^n from a
                  periodic finite difference written here with np.roll;
  inc pressure    without viscosity the z-solve is the identity, so one
                  incremental step must give the non-incremental velocity and
                  the non-incremental pressure up to a constant (a wrong sign
                  of the added -grad p_ext breaks the pressure check); the
                  pressure history after the bootstrap and after each step.
Plus the reference's option checks (solver.py:274-281).
"""

import math

import numpy as np
import pytest

from oracle import oracle as orc

NX, NY, NZ = 16, 12, 16


def _oracle(Re=2000.0, seed=5, **kw):
    co = orc.ChannelOracle(NX, NY, NZ, 2 * math.pi, math.pi, 2.0, 1.5, Re, order=4, **kw)
    rng = np.random.default_rng(seed)
    zc = co.z_c
    shapes = ((NX + 1, NY + 2, NZ + 2), (NX + 2, NY + 1, NZ + 2), (NX + 2, NY + 2, NZ + 1))
    f = [0.05 * rng.standard_normal(s) for s in shapes]
    f[0] += (1.5 * (1.0 - (zc - 1.0) ** 2)).reshape(1, 1, -1)  # Poiseuille
    co.set_fields(*f, np.zeros((NX + 2, NY + 2, NZ + 2)))
    co.apply_bc_uvw()
    co.project_initial()
    return co


def _clone(co, **kw):
    c = _oracle(**kw)
    c.set_fields(co.u, co.v, co.w, co.p)
    return c


def _lap_xy_u(u, nu, dx, dy):
    """nu (d2x + d2y) of u's interior, periodic in x and y (np.roll)."""
    U = u[1 : NX + 1, 1 : NY + 1, 1 : NZ + 1]
    d2x = (np.roll(U, -1, 0) - 2.0 * U + np.roll(U, 1, 0)) / dx**2
    d2y = (np.roll(U, -1, 1) - 2.0 * U + np.roll(U, 1, 1)) / dy**2
    return nu * (d2x + d2y)


def test_option_checks(check):
    for kw, msg in (({"bdf2_pressure": "incremental"}, "bdf2_pressure"),
                    ({"bdf2_xy_rhs": "ab2"}, "bdf2_xy_rhs")):
        with pytest.raises(ValueError, match=msg):
            orc.ChannelOracle(NX, NY, NZ, 1.0, 1.0, 2.0, 1.5, 100.0, **kw)
    check("invalid option strings rejected", True, "")


def test_lagged_equals_extrap_for_a_current_history(check):
    dt = 0.01
    ex = _oracle()
    lg = _clone(ex, bdf2_xy_rhs="lagged")
    for c in (ex, lg):
        c.set_history(c.u, c.v, c.w, dt)
    a, b = ex.bdf2_predictor(dt), lg.bdf2_predictor(dt)
    same = all(np.array_equal(x, y) for x, y in zip(a[:3], b[:3]))
    check("V^{n-1} = V^n: lagged predictor == extrap predictor", same,
          f"max|d|={max(np.abs(x - y).max() for x, y in zip(a[:3], b[:3])):.1e}")


def test_lagged_differs_by_dt_eff_R(check):
    dt = 0.01
    ex = _oracle(seed=9)
    lg = _clone(ex, bdf2_xy_rhs="lagged")
    zero = [np.zeros_like(x) for x in (ex.u, ex.v, ex.w)]
    for c in (ex, lg):
        c.set_history(*zero, dt)  # R^{n-1} = 0: extrap Rh = 2 R^n, lagged R^n
    a, b = ex.bdf2_predictor(dt), lg.bdf2_predictor(dt)
    dt_eff = 2.0 / 3.0 * dt
    d = (a[0] - b[0])[1 : NX + 1, 1 : NY + 1, 1 : NZ + 1]
    R = _lap_xy_u(ex.u, ex.nu, ex.dx, ex.dy)
    err = np.abs(d - dt_eff * R).max()
    scale = np.abs(a[0]).max()
    check("u: extrap - lagged == dt_eff nu L_xy u (interior)", err <= 1e-14 * scale,
          f"max err {err:.1e}, |pred| {scale:.2f}, |dt_eff R| {np.abs(dt_eff * R).max():.1e}")
    outside = np.ones(a[0].shape, dtype=bool)
    outside[1 : NX + 1, 1 : NY + 1, 1 : NZ + 1] = False
    ghost = np.abs((a[0] - b[0])[outside]).max()
    check("u: identical outside the interior", ghost == 0.0, f"{ghost:.1e}")


def test_inc_without_viscosity_equals_noninc(check):
    dt = 0.02
    non = _oracle(Re=1e300, seed=3)
    inc = _clone(non, Re=1e300, bdf2_pressure="inc")
    for n in range(4):
        non.step_sl_bdf2(dt)
        inc.step_sl_bdf2(dt)
        du = max(np.abs(a - b).max() for a, b in ((non.u, inc.u), (non.v, inc.v),
                                                  (non.w, inc.w)))
        dp = (inc.p - non.p)[1 : NX + 1, 1 : NY + 1, 1 : NZ + 1]
        spread = dp.max() - dp.min()
        pscale = np.abs(non.p).max()
        check(f"step {n + 1}: velocity inc == noninc", du <= 1e-12, f"max|du|={du:.1e}")
        check(f"step {n + 1}: pressure inc == noninc + const", spread <= 1e-11 * pscale,
              f"spread {spread:.1e} of |p| {pscale:.1e}")


def test_inc_pressure_history(check):
    dt = 0.01
    co = _oracle(seed=4, bdf2_pressure="inc")
    check("no extrapolation before any step", co.pressure_extrapolation() is None, "")
    co.step_sl_bdf2(dt)  # BDF1 bootstrap
    p1 = co.p.copy()
    check("after the bootstrap p_ext = p^1", np.array_equal(co.pressure_extrapolation(), p1), "")
    co.step_sl_bdf2(dt)
    p2 = co.p.copy()
    pe = co.pressure_extrapolation()
    check("after one BDF2 step p_ext = 2 p^2 - p^1", np.array_equal(pe, 2.0 * p2 - p1), "")
