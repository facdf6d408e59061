"""Pins the statistics oracle (oracle/turbstats.py) by closed forms.

The constructed case is the one of the reference's
tests/test_stats_plane_mean.py (a fixed zero-mean face pattern plus a plane-
uniform offset alternating +a, -a, so var_t(U) = a^2 exactly).  The other
cases are separable patterns and single Fourier modes whose statistics are
known in closed form; each is chosen so that a plausible slip fails it:
cell-centre instead of node variance (the cos(theta/2) filter factor), a
missing time-mean correction, a transposed or unfolded spectrum, a wrong
normalisation, a lost top-wall sign.
"""

import math

import numpy as np

from oracle.turbstats import StatsOracle, fold, snapshot


def _zeros(nx, ny, nz):
    return (np.zeros((nx + 1, ny + 2, nz + 2)), np.zeros((nx + 2, ny + 1, nz + 2)),
            np.zeros((nx + 2, ny + 2, nz + 1)))


def test_time_mean_correction_reference_case(check):
    nx, ny, nz = 16, 12, 24
    i = np.arange(nx + 1, dtype=np.float64)
    face = np.cos(2 * math.pi * i / nx)
    pattern_var = float((face[1:] ** 2).mean())
    cc = 0.5 * (face[:-1] + face[1:])
    pattern_var_filtered = float((cc ** 2).mean())
    u, v, w = _zeros(nx, ny, nz)
    st = StatsOracle(nx, ny, nz)
    a = 0.37
    for n in range(8):
        u[:] = (face + (a if n % 2 == 0 else -a))[:, None, None]
        st.accumulate(u, v, w)
    raw = st.prof[1, nz // 2] / st.n_samples
    out = st.finalize(1e-3, 0.01)
    got = out["uu_mean"][nz // 2]
    check("corrected u'u' = pattern_var + a^2", abs(got - (pattern_var + a * a)) < 1e-12,
          f"{got!r} vs {pattern_var + a * a!r}")
    check("raw accumulator = pattern_var (biased)", abs(raw - pattern_var) < 1e-12, f"{raw!r}")
    check("u'u' at the u nodes, not cell centres",
          abs(raw - pattern_var_filtered) > 0.5 * (pattern_var - pattern_var_filtered),
          f"raw {raw:.6f} filtered {pattern_var_filtered:.6f}")
    check("<u'w'> zero", abs(out["uw_mean"][nz // 2]) < 1e-14, f"{out['uw_mean'][nz // 2]:.1e}")
    check("w'w' zero", abs(out["ww_mean"][nz // 2]) < 1e-14, f"{out['ww_mean'][nz // 2]:.1e}")
    check("U time mean zero", abs(out["U_mean"][nz // 2]) < 1e-15, f"{out['U_mean'][nz // 2]:.1e}")


def test_separable_profiles_closed_form(check):
    nx, ny, nz = 16, 12, 10
    mx, my = 3, 2
    tx, ty = 2 * math.pi * mx / nx, 2 * math.pi * my / ny
    rng = np.random.default_rng(5)
    A = rng.uniform(0.5, 1.5, nz)       # u amplitude per level
    B = rng.uniform(0.5, 1.5, nz)       # v amplitude per level
    C = rng.uniform(0.5, 1.5, nz + 1)   # w amplitude per face
    U0, V0 = rng.uniform(-1, 1, nz), rng.uniform(-1, 1, nz)
    u, v, w = _zeros(nx, ny, nz)
    i = np.arange(nx + 1)
    # u_int[i] at face i; v_int[:, j] at y-face j; w_int[i] at x-centre i + 1/2
    u[:, 1:ny + 1, 1:nz + 1] = (U0[None, None, :] + A[None, None, :]
                                * np.cos(tx * i)[:, None, None])
    j = np.arange(ny + 1)
    v[1:nx + 1, :, 1:nz + 1] = (V0[None, None, :] + B[None, None, :]
                                * np.cos(ty * j)[None, :, None])
    xc = np.arange(nx) + 0.5
    w[1:nx + 1, 1:ny + 1, :] = C[None, None, :] * np.cos(tx * xc)[:, None, None]
    s = snapshot(u, v, w, nx, ny, nz)
    Cc = 0.5 * (C[:-1] + C[1:])

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))

    check("U = U0 (cell-centred mean)", rel(s["U"], U0) < 1e-14, f"{rel(s['U'], U0):.1e}")
    check("V = V0", rel(s["V"], V0) < 1e-14, f"{rel(s['V'], V0):.1e}")
    check("W = 0", np.max(np.abs(s["W"])) < 1e-15, f"{np.max(np.abs(s['W'])):.1e}")
    check("u'u' = A^2/2 at nodes", rel(s["uu"], A ** 2 / 2) < 1e-14, f"{rel(s['uu'], A**2/2):.1e}")
    check("v'v' = B^2/2 at nodes", rel(s["vv"], B ** 2 / 2) < 1e-14, f"{rel(s['vv'], B**2/2):.1e}")
    ww = 0.5 * (C[:-1] ** 2 / 2 + C[1:] ** 2 / 2)
    check("w'w' = face variances averaged to centres", rel(s["ww"], ww) < 1e-14,
          f"{rel(s['ww'], ww):.1e}")
    uw = A * math.cos(tx / 2) * Cc / 2
    check("<u'w'> = A cos(theta/2) Cc / 2 (cell-centred pair)", rel(s["uw"], uw) < 1e-14,
          f"{rel(s['uw'], uw):.1e}")


def _mode_plane(nx, ny, modes):
    i = np.arange(nx)[:, None]
    j = np.arange(ny)[None, :]
    out = np.zeros((nx, ny))
    for kx, ky, amp, ph in modes:
        out += amp * np.cos(2 * math.pi * (kx * i / nx + ky * j / ny) + ph)
    return out


def test_spectra_modes_fold_and_signs(check):
    nx, ny, nz = 16, 12, 8
    kb, kt = 2, 7
    rng = np.random.default_rng(9)
    # distinct folded bins, 0 < |kx| < nx/2, 0 < ky < ny/2; negative kx folds onto |kx|
    modes = [(3, 2, 0.8, 0.3), (-2, 1, 0.5, 1.1), (5, 4, 0.3, -0.7), (1, 3, 0.6, 2.0)]
    expect_uu = np.zeros((nx // 2, ny // 2))
    for kx, ky, amp, _ in modes:
        expect_uu[abs(kx) - 1, ky - 1] += amp * amp / 4
    du = 0.4  # phase offset of w against u
    wmodes = [(kx, ky, amp, ph + du) for kx, ky, amp, ph in modes]
    pu = _mode_plane(nx, ny, modes)
    pw = _mode_plane(nx, ny, wmodes)
    u, v, w = _zeros(nx, ny, nz)
    u[1:nx + 1, 1:ny + 1, kb] = pu + 0.25           # u_node at level kb (plus a plane mean)
    v[1:nx + 1, 1:ny + 1, kt] = pu                  # v_node at level kt only
    w[1:nx + 1, 1:ny + 1, kb - 1] = pw
    w[1:nx + 1, 1:ny + 1, kb] = pw                  # cell-centred w at level kb = pw
    u[1:nx + 1, 1:ny + 1, kt] = pu
    w[1:nx + 1, 1:ny + 1, kt - 1] = pu
    w[1:nx + 1, 1:ny + 1, kt] = pu                  # top: w = u, negated by the top-wall rule
    st = StatsOracle(nx, ny, nz, planes=[(kb, kt)])
    st.accumulate(u, v, w)
    E = st.spec[0]
    tol = 1e-14
    # uu: both planes carry pu -> average = expect_uu
    check("E_uu at the modes' folded bins", np.max(np.abs(E[0] - expect_uu)) < tol,
          f"{np.max(np.abs(E[0] - expect_uu)):.1e}")
    check("E_vv: only the top plane -> half", np.max(np.abs(E[1] - expect_uu / 2)) < tol,
          f"{np.max(np.abs(E[1] - expect_uu / 2)):.1e}")
    check("E_ww: bottom pw, top pu -> same amplitudes", np.max(np.abs(E[2] - expect_uu)) < tol,
          f"{np.max(np.abs(E[2] - expect_uu)):.1e}")
    # uw: bottom cos(du) * E, top -E (w negated) -> 0.5 (cos du - 1) E
    exp_uw = 0.5 * (math.cos(du) - 1.0) * expect_uu
    check("E_uw: phase cosine at the bottom, sign flip at the top",
          np.max(np.abs(E[3] - exp_uw)) < tol, f"{np.max(np.abs(E[3] - exp_uw)):.1e}")
    # the folded half-plane of a field with no kx=0 / ky=0 / Nyquist content
    # holds half its variance (Parseval)
    var = float(np.mean((pu - pu.mean()) ** 2))
    check("Parseval over the folded half-plane", abs(E[0].sum() - 0.5 * var) < 1e-14,
          f"{E[0].sum()!r} vs {0.5 * var!r}")
    # fold() itself: E[a+1, b+1] + E[nx-1-a, b+1]
    R = rng.standard_normal((nx, ny // 2 + 1))
    F = fold(R, nx, ny)
    a, b = 2, 3
    check("fold index map", F[a, b] == R[a + 1, b + 1] + R[nx - 1 - a, b + 1], "")


def test_finalize_u_tau(check):
    nx, ny, nz = 8, 6, 6
    u, v, w = _zeros(nx, ny, nz)
    prof = np.linspace(0.2, 0.9, nz)
    u[:, 1:ny + 1, 1:nz + 1] = prof[None, None, :]
    st = StatsOracle(nx, ny, nz)
    st.accumulate(u, v, w)
    nu, zc1 = 2e-3, 0.05
    d = st.finalize(nu, zc1, "dirichlet")
    n = st.finalize(nu, zc1, "neumann")
    check("u_tau both walls", abs(d["u_tau"] - math.sqrt(nu * 0.5 * (0.2 + 0.9) / zc1)) < 1e-15,
          f"{d['u_tau']!r}")
    check("u_tau bottom wall only (free-slip top)",
          abs(n["u_tau"] - math.sqrt(nu * 0.2 / zc1)) < 1e-15, f"{n['u_tau']!r}")
    check("no samples raises", _raises(lambda: StatsOracle(4, 4, 4).finalize(1.0, 1.0)), "")


def _raises(fn):
    try:
        fn()
    except RuntimeError:
        return True
    return False
