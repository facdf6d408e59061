"""Pin the oracle against the golden vectors written from the reference.

The fixtures (tests/golden/*.npz, made by tests/golden/make_golden.py) hold the
reference's own outputs on seeded inputs.  These tests run on the CPU only and
are what makes the oracle a trustworthy checker for the HIP layer.

Tolerances: the grid and tables differ from the reference only through
libm/SLEEF tanh rounding (a few ulp); feet and stencil indices must be
identical given identical tables (ties -- a z exactly on a face -- excepted,
see `tie_band`); advected values within 1e-13 of the max; solves and full
steps within 1e-12 relative (PCR/Thomas and pocketfft/MKL rounding).
"""

import math

import numpy as np
import pytest

from oracle import oracle as orc


def rel_err(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)


@pytest.mark.parametrize("name", ["sym_g15_n16", "bot_g18_n24", "sym_g15_n64"])
def test_grid_and_tables(check, golden, name):
    g = golden("grid_" + name)
    gamma, nz, Lz = float(g["gamma"]), int(g["nz"]), float(g["Lz"])
    st = str(g["stretching"])
    z_f, z_c, dz_f, dz_c = orc.generate_grid(gamma, nz, Lz, st)
    check("z_f vs reference", np.abs(z_f - g["z_f"]).max() < 4e-16 * Lz,
          f"max|dz_f|={np.abs(z_f - g['z_f']).max():.1e}")
    check("z_c vs reference", np.abs(z_c - g["z_c"]).max() < 4e-16 * Lz,
          f"max|dz_c|={np.abs(z_c - g['z_c']).max():.1e}")
    # the tables from the reference's own grid must be identical to ulp level
    for comp, nodes in (("u", g["z_c"]), ("v", g["z_c"]), ("w", g["z_f"])):
        for o in (2, 4, 6):
            t = orc.z_denominator_table(nodes, o)
            e = rel_err(t, g[f"dinv_{comp}{o}"])
            check(f"dinv {comp} order {o}", e < 1e-15, f"rel={e:.1e}")
    # the analytic inverse map (tanh(gamma) as the reference computes it)
    sl = orc.SLOracle(8, 6, nz, 0.5, 0.5, 4.0, 3.0, Lz, g["z_f"], g["z_c"], gamma,
                      stretching_type=st, order=6)
    sl.set_tanh_g(float(g["tanh_g"]))
    check("tanh(gamma)", math.tanh(gamma) == float(g["tanh_g"]) or
          abs(math.tanh(gamma) - float(g["tanh_g"])) < 2.3e-16, "")
    check("z_lo/z_hi", sl.z_lo == float(g["z_lo"]) and sl.z_hi == float(g["z_hi"]),
          f"{sl.z_lo} {sl.z_hi}")
    kf = sl.face_coord(g["zq"])
    err = np.abs(kf - g["kf"]).max()
    check("face_coord vs reference", err < 1e-12, f"max|dkf|={err:.1e}")
    # floor(kf) may differ only where kf sits within a few ulp of an integer
    fl_o, fl_r = np.floor(kf), np.floor(g["kf"])
    tie = np.abs(g["kf"] - np.round(g["kf"])) < 1e-12
    check("floor(kf) identical off ties", np.all(fl_o[~tie] == fl_r[~tie]),
          f"ties={int(tie.sum())} mismatches_off_tie={int((fl_o[~tie] != fl_r[~tie]).sum())}")


def _sl_from(g):
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    Lx, Ly, Lz, gamma = float(g["Lx"]), float(g["Ly"]), float(g["Lz"]), float(g["gamma"])
    return orc.SLOracle(nx, ny, nz, Lx / nx, Ly / ny, Lx, Ly, Lz, g["z_f"], g["z_c"], gamma,
                        stretching_type=str(g["stretching"]), order=int(g["order"]),
                        traj_order=int(g["traj_order"]) if "traj_order" in g else 2)


@pytest.mark.parametrize("name", ["o4", "o6", "o6_bottom", "o4_clamp", "o6_traj4"])
def test_advect_vs_reference(check, golden, name):
    g = golden("advect_" + name)
    sl = _sl_from(g)
    dt = float(g["dt"])
    for comp in "uvw":
        xd, yd, zd = sl.departure(comp, g["um"], g["vm"], g["wm"], dt)
        same = (xd == g[f"xd_{comp}"]) & (yd == g[f"yd_{comp}"]) & (zd == g[f"zd_{comp}"])
        dev = max(np.abs(xd - g[f"xd_{comp}"]).max(), np.abs(yd - g[f"yd_{comp}"]).max(),
                  np.abs(zd - g[f"zd_{comp}"]).max())
        # A w arrival sits exactly on a face node; when its trajectory sample
        # stays on that face, glibc's and SLEEF's atanh may pick the two
        # neighbouring intervals (a tie), giving x*(1/x) weights that differ by
        # one ulp.  Those feet may differ by a few ulp; all others are identical.
        check(f"feet {comp} bit-exact off ties",
              same.mean() >= 0.995 and dev <= 4 * np.finfo(float).eps * float(g["Lx"]),
              f"identical={int(same.sum())}/{same.size} max|d|={dev:.1e}")
        field = {"u": g["u"], "v": g["v"], "w": g["w"]}[comp]
        vals, _, (ix, iy, k0) = sl.gather(comp, field, g[f"xd_{comp}"], g[f"yd_{comp}"],
                                          g[f"zd_{comp}"], want_index=True)
        idx_ok = (ix == g[f"ix_{comp}"]) & (iy == g[f"iy_{comp}"]) & (k0 == g[f"k0_{comp}"])
        check(f"stencil indices {comp} bit-exact", idx_ok.all(),
              f"identical={int(idx_ok.sum())}/{idx_ok.size}")
    us, vs, ws = sl.advect(g["u"], g["v"], g["w"], g["um"], g["vm"], g["wm"], dt)
    for nm, a, b in (("u", us, g["ustar"]), ("v", vs, g["vstar"]), ("w", ws, g["wstar"])):
        e = rel_err(a, b)
        check(f"advect {nm}", e < 1e-13, f"rel={e:.1e}")
    check("n_clamped", sl.n_clamped_last == int(g["n_clamped"]),
          f"{sl.n_clamped_last} vs {int(g['n_clamped'])}")


@pytest.mark.parametrize("tag", ["sym", "bot"])
def test_solves_vs_reference(check, golden, tag):
    g = golden("solve_small")
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    Lx, Ly = float(g["Lx"]), float(g["Ly"])
    dx, dy = Lx / nx, Ly / ny
    z_f, z_c = g[f"{tag}_z_f"], g[f"{tag}_z_c"]
    dz_f, dz_c = z_f[1:] - z_f[:-1], z_c[1:] - z_c[:-1]
    top = str(g[f"{tag}_top"])
    fft = orc.initialize_fft_solver(nx, ny, nz, dx, dy, dz_c, dz_f, top)
    p = orc.solve_poisson_fft(g[f"{tag}_div"], fft)
    e = rel_err(p, g[f"{tag}_p"])
    check("poisson p", e < 1e-12, f"rel={e:.1e}")
    nu, dt = float(g[f"{tag}_nu"]), float(g[f"{tag}_dt"])
    for theta, th in ((1.0, "t1"), (0.5, "t05")):
        su = orc.solve_implicit_diffusion_u(g[f"{tag}_u"], dt, nx, ny, nz, dz_c, dz_f, nu,
                                            theta, top)
        sv = orc.solve_implicit_diffusion_v(g[f"{tag}_v"], dt, nx, ny, nz, dz_c, dz_f, nu,
                                            theta, top)
        sw = orc.solve_implicit_diffusion_w(g[f"{tag}_w"], dt, nx, ny, nz, dz_c, dz_f, nu,
                                            theta)
        for nm, a, b in (("u", su, g[f"{tag}_{th}_u"]), ("v", sv, g[f"{tag}_{th}_v"]),
                         ("w", sw, g[f"{tag}_{th}_w"])):
            e = rel_err(a, b)
            check(f"zsolve {nm} theta={theta}", e < 1e-13, f"rel={e:.1e}")


def _bdf2_options(g):
    """The fixture's BDF2 options (the defaults for files without them)."""
    opts = {k: str(g[k]) if k in g else d
            for k, d in (("bdf2_pressure", "noninc"), ("bdf2_xy_rhs", "extrap"))}
    opts["traj_order"] = int(g["traj_order"]) if "traj_order" in g else 2
    return opts


@pytest.mark.parametrize("name", ["o6", "o4_bottom_neumann", "o4_inc", "o6_lagged", "o4_traj4"])
def test_step_vs_reference(check, golden, name):
    g = golden("step_" + name)
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    co = orc.ChannelOracle(nx, ny, nz, float(g["Lx"]), float(g["Ly"]), float(g["Lz"]),
                           float(g["gamma"]), float(g["Re"]), order=int(g["order"]),
                           stretching_type=str(g["stretching"]),
                           top_wall_bc_type=str(g["top"]), z_f=g["z_f"], z_c=g["z_c"],
                           **_bdf2_options(g))
    co.set_fields(g["u0"], g["v0"], g["w0"], g["p0"])
    dt = float(g["dt"])
    n_steps = int(g["n_steps"])
    for n in range(n_steps):
        ub, fo = co.step_sl_bdf2(dt)
        check(f"step {n + 1} u_bulk", abs(ub - g["u_bulk"][n]) < 1e-14, f"{ub!r}")
        check(f"step {n + 1} forcing", abs(fo - g["forcing"][n]) < 1e-12 * max(1.0, abs(fo)),
              f"{fo:.6e} vs {g['forcing'][n]:.6e}")
        check(f"step {n + 1} n_clamped", co.n_clamped_last == g["n_clamped"][n], "")
        if f"u_{n + 1}" in g:
            for nm, a in (("u", co.u), ("v", co.v), ("w", co.w), ("p", co.p)):
                e = rel_err(a, g[f"{nm}_{n + 1}"])
                check(f"step {n + 1} {nm}", e < 1e-12, f"rel={e:.1e}")


@pytest.mark.parametrize("tag", ["sym", "bot"])
def test_imex_rhs_vs_reference(check, golden, tag):
    """compute_momentum_rhs_fused_imex (operators.py:639-835): the oracle
    restates the reference's element operations in order, so bit-identical."""
    g = golden("imex_small")
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    z_f, z_c = g[f"{tag}_z_f"], g[f"{tag}_z_c"]
    dz_f, dz_c = z_f[1:] - z_f[:-1], z_c[1:] - z_c[:-1]
    dx, dy = float(g["Lx"]) / nx, float(g["Ly"]) / ny
    out = orc.compute_momentum_rhs_fused_imex(g[f"{tag}_u"], g[f"{tag}_v"], g[f"{tag}_w"],
                                              nx, ny, nz, dx, dy, dz_c, dz_f,
                                              float(g[f"{tag}_nu"]))
    for name, a in zip(("ru", "rv", "rw"), out):
        b = g[f"{tag}_{name}"]
        check(f"{tag} {name} bit-identical to the reference", np.array_equal(a, b),
              f"max|d|={np.abs(a - b).max():.1e} of {np.abs(b).max():.1e}")


def test_imex_steps_vs_reference(check, golden):
    """SLChannelFlow.step_imex (solver.py:1055-1134): three steps, forward
    Euler then AB2, Crank-Nicolson z-solves, projection, bulk shift."""
    g = golden("imex_small")
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    co = orc.ChannelOracle(nx, ny, nz, float(g["Lx"]), float(g["Ly"]), float(g["Lz"]),
                           float(g["gamma"]), float(g["step_Re"]), order=6,
                           z_f=g["step_z_f"], z_c=g["step_z_c"])
    co.set_fields(g["step_u0"], g["step_v0"], g["step_w0"], g["step_p0"])
    dt = float(g["step_dt"])
    for n in range(int(g["step_n"])):
        ub, fo = co.step_imex(dt)
        check(f"step {n + 1} u_bulk", abs(ub - g["step_u_bulk"][n]) < 1e-14, f"{ub!r}")
        check(f"step {n + 1} forcing",
              abs(fo - g["step_forcing"][n]) < 1e-12 * max(1.0, abs(fo)),
              f"{fo:.6e} vs {g['step_forcing'][n]:.6e}")
        for nm in "uvwp":
            e = rel_err(getattr(co, nm), g[f"step_{nm}_{n + 1}"])
            check(f"step {n + 1} {nm}", e < 1e-12, f"rel={e:.1e}")


def test_stats_vs_reference(check, golden):
    """TurbulenceStats.accumulate_statistics over three snapshots with two
    extra spectra planes, and finalize_statistics (turbstats.py:156-467)."""
    from oracle.turbstats import PROFILE_NAMES, StatsOracle

    g = golden("stats_small")
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    planes = [(int(g["k_bot"]), int(g["k_top"]))] + [tuple(int(x) for x in r)
                                                      for r in g["extra_k"]]
    st = StatsOracle(nx, ny, nz, planes=planes)
    for n in range(3):
        st.accumulate(g[f"u{n}"], g[f"v{n}"], g[f"w{n}"])
    for i, name in enumerate(PROFILE_NAMES):
        e = rel_err(st.prof[i], g[name])
        check(f"{name} within 1e-14 of the reference", e < 1e-14, f"rel={e:.1e}")
    for c, comp in enumerate(("uu", "vv", "ww", "uw")):
        e = rel_err(st.spec[0][c], g[f"E_{comp}_2d_sum"])
        check(f"E_{comp}_2d_sum within 1e-14", e < 1e-14, f"rel={e:.1e}")
        for p in range(len(planes) - 1):
            e = rel_err(st.spec[p + 1][c], g[f"extra{p}_{comp}"])
            check(f"extra plane {p} {comp} within 1e-14", e < 1e-14, f"rel={e:.1e}")
    fin = st.finalize(float(g["nu"]), float(g["z_c"][1]))
    e = rel_err(fin["U_mean"], g["fin_U_mean"])
    check("finalized U_mean within 1e-14", e < 1e-14, f"rel={e:.1e}")
    # the stresses are sum/n + (squared-mean sum)/n - mean^2: a cancelling
    # difference, so their rounding is measured on the scale of the squared
    # plane means (the largest second moment), as the plane-mean sums of the
    # two libraries differ in their last bits
    scale2 = float(np.abs(g["Um2_sum"]).max()) / 3.0
    for k in ("uu_mean", "vv_mean", "ww_mean", "uw_mean"):
        e = float(np.abs(fin[k] - g[f"fin_{k}"]).max()) / scale2
        check(f"finalized {k} within 1e-14 of the second-moment scale", e < 1e-14,
              f"rel={e:.1e} (scale {scale2:.2f})")
    for c, comp in enumerate(("uu", "vv", "ww", "uw")):
        e = rel_err(fin["spectra"][0][c], g[f"fin_E_{comp}_2d"])
        check(f"finalized E_{comp}_2d within 1e-14", e < 1e-14, f"rel={e:.1e}")
    check("u_tau within 1e-15", abs(fin["u_tau"] - float(g["fin_u_tau"]))
          < 1e-15 * float(g["fin_u_tau"]),
          f"{fin['u_tau']!r} vs {float(g['fin_u_tau'])!r}")


def test_diagnostics_vs_reference(check, golden):
    """compute_u_tau (utils.py:116-146) both top-wall types and
    compute_cfl_fused (operators.py:839-940)."""
    g = golden("diag_small")
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    z_f, z_c = g["z_f"], g["z_c"]
    dz_f, dz_c = z_f[1:] - z_f[:-1], z_c[1:] - z_c[:-1]
    dx, dy = float(g["Lx"]) / nx, float(g["Ly"]) / ny
    nu = float(g["nu"])
    for top in ("dirichlet", "neumann"):
        ut = orc.compute_u_tau(g["u"], z_c, nu, top)
        ref = float(g[f"u_tau_{top}"])
        # the plane mean of the first level sums in numpy's order, not torch's
        check(f"u_tau ({top}) within 1e-15", abs(ut - ref) <= 1e-15 * ref,
              f"{ut!r} vs {ref!r}")
    dti = orc.compute_cfl_fused(g["u"], g["v"], g["w"], nx, ny, nz, dx, dy, dz_f, dz_c)
    check("compute_cfl_fused bit-identical", dti == float(g["dti"]),
          f"{dti!r} vs {float(g['dti'])!r}")
