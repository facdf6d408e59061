"""Parity of the HIP layer (libslc_hip.so) against the reference's golden
vectors and against the CPU oracle, on identical inputs.

Bars (stated per check): feet bit-exact except z-face ties; stencil indices
bit-exact; ghost fill, divergence and projection bit-exact; advected fields
within 1e-13 of the max; solves and full steps within 1e-12 relative
(Thomas vs PCR, rocFFT vs pocketfft rounding); max|div| < 1e-11 and the bulk
velocity pinned at 1.0 after every step.
"""

import math

import numpy as np
import pytest
import torch

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

EPS = np.finfo(float).eps


def rel_err(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = b.detach().cpu().numpy() if isinstance(b, torch.Tensor) else np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def cuda(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a)).to(dtype).cuda()


def make_plan(g, order=None, nu=0.0, top=None):
    from slchannel_amd import HipPlan

    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    return HipPlan(nx, ny, nz, float(g["Lx"]), float(g["Ly"]), float(g["Lz"]), g["z_f"],
                   g["z_c"], float(g["gamma"]),
                   stretching_type=str(g.get("stretching", "symmetric")),
                   order=int(order if order is not None else g["order"]),
                   traj_order=int(g["traj_order"]) if "traj_order" in g else 2,
                   top_wall_bc_type=top or str(g.get("top", "dirichlet")), nu=nu)


def feet_ok(hip, ref, Lx):
    same = np.all([(h == r) for h, r in zip(hip, ref)], axis=0)
    dev = max(float(np.abs(h - r).max()) for h, r in zip(hip, ref))
    return same.mean() >= 0.995 and dev <= 4 * EPS * Lx, same, dev


# ---------------------------------------------------------------------------


@pytest.mark.parametrize("top", ["dirichlet", "neumann"])
def test_ghost_fill_bitexact(check, top):
    from slchannel_amd import HipPlan
    from slchannel_amd.grid import generate_grid

    nx, ny, nz = 13, 11, 17
    z_f, z_c, _, _ = generate_grid(1.5, nz, 2.0)
    plan = HipPlan(nx, ny, nz, 4.0, 3.0, 2.0, z_f, z_c, 1.5, order=4, top_wall_bc_type=top)
    rng = np.random.default_rng(0)
    u, v, w = (rng.normal(size=s) for s in ((nx + 1, ny + 2, nz + 2), (nx + 2, ny + 1, nz + 2),
                                             (nx + 2, ny + 2, nz + 1)))
    tu, tv, tw = cuda(u), cuda(v), cuda(w)
    plan.apply_bc(tu, tv, tw)
    orc.apply_bc_all(u, v, w, top)
    for nm, a, b in (("u", tu, u), ("v", tv, v), ("w", tw, w)):
        check(f"bc {nm} ({top})", np.array_equal(a.cpu().numpy(), b), "bit-exact")


@pytest.mark.parametrize("name", ["o4", "o6", "o6_bottom", "o4_clamp", "o6_traj4"])
def test_feet_stencils_advect_vs_reference(check, golden, name):
    g = golden("advect_" + name)
    plan = make_plan(g)
    Lx = float(g["Lx"])
    dt = float(g["dt"])
    um, vm, wm = cuda(g["um"]), cuda(g["vm"]), cuda(g["wm"])
    for comp in "uvw":
        n = plan.n_arrival(comp)
        xd, yd, zd = (torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(3))
        plan.departure(comp, um, vm, wm, dt, xd, yd, zd)
        ok, same, dev = feet_ok([t.cpu().numpy() for t in (xd, yd, zd)],
                                [g[f"{c}d_{comp}"] for c in "xyz"], Lx)
        check(f"feet {comp} vs reference", ok,
              f"identical={int(same.sum())}/{same.size} max|d|={dev:.1e}")
        ix, iy, k0 = (torch.empty(n, dtype=torch.int32, device="cuda") for _ in range(3))
        gx, gy, gz = (cuda(g[f"{c}d_{comp}"]) for c in "xyz")
        plan.stencil(comp, gx, gy, gz, ix, iy, k0)
        same_idx = ((ix.cpu().numpy() == g[f"ix_{comp}"]) & (iy.cpu().numpy() == g[f"iy_{comp}"])
                    & (k0.cpu().numpy() == g[f"k0_{comp}"]))
        check(f"stencil indices {comp} bit-exact", same_idx.all(),
              f"identical={int(same_idx.sum())}/{same_idx.size}")
    from slchannel_amd import SLAdvector

    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    adv = SLAdvector(nx, ny, nz, Lx / nx, float(g["Ly"]) / ny, Lx, float(g["Ly"]),
                     float(g["Lz"]), g["z_f"], g["z_c"], float(g["gamma"]),
                     stretching_type=str(g["stretching"]), order=int(g["order"]), plan=plan)
    us, vs, ws = adv.advect(cuda(g["u"]), cuda(g["v"]), cuda(g["w"]), um, vm, wm, dt)
    for nm, a, b in (("u", us, g["ustar"]), ("v", vs, g["vstar"]), ("w", ws, g["wstar"])):
        e = rel_err(a, b)
        check(f"advect {nm} vs reference", e < 1e-13, f"rel={e:.1e}")
    ncl = int(adv.n_clamped_last.item())
    check("n_clamped", ncl == int(g["n_clamped"]), f"{ncl} vs {int(g['n_clamped'])}")


def _random_case(nx, ny, nz, order, stretching, seed, gamma=1.6, cfl=2.5):
    from slchannel_amd.grid import generate_grid

    Lx, Ly, Lz = 4 * math.pi, 2 * math.pi, 2.0
    z_f, z_c, _, _ = (t.numpy() for t in generate_grid(gamma, nz, Lz, stretching_type=stretching))
    rng = np.random.default_rng(seed)
    shapes = ((nx + 1, ny + 2, nz + 2), (nx + 2, ny + 1, nz + 2), (nx + 2, ny + 2, nz + 1))
    # smooth fields from a few random modes plus noise-free ghosts
    fields = []
    for comp, s in zip("uvw", shapes):
        I = np.arange(s[0])[:, None, None]
        J = np.arange(s[1])[None, :, None]
        zz = (z_c if comp != "w" else z_f)[None, None, :]
        f = np.zeros(s)
        for _ in range(4):
            a = rng.normal()
            f += a * np.sin(2 * math.pi * rng.integers(1, 3) * I / nx + rng.uniform(0, 6)) * \
                np.cos(2 * math.pi * rng.integers(0, 3) * J / ny + rng.uniform(0, 6)) * \
                np.cos(math.pi * rng.uniform(0.5, 2) * zz / Lz + rng.uniform(0, 6))
        fields.append(f)
    mids = [1.0 + 0.3 * fields[0], 0.3 * fields[1], 0.2 * fields[2]]
    for arr in (fields, mids):
        orc.apply_bc_all(*arr)
    dx = Lx / nx
    dt = cfl * dx / float(np.abs(mids[0]).max())
    return dict(nx=nx, ny=ny, nz=nz, Lx=Lx, Ly=Ly, Lz=Lz, gamma=gamma, order=order,
                stretching=stretching, z_f=z_f, z_c=z_c, dt=dt, fields=fields, mids=mids)


@pytest.mark.parametrize("order,stretching", [(4, "symmetric"), (6, "symmetric"),
                                              (6, "bottom")])
def test_advect_vs_oracle(check, order, stretching):
    c = _random_case(40, 36, 48, order, stretching, seed=order + len(stretching))
    plan = make_plan(c)
    sl = orc.SLOracle(c["nx"], c["ny"], c["nz"], c["Lx"] / c["nx"], c["Ly"] / c["ny"], c["Lx"],
                      c["Ly"], c["Lz"], c["z_f"], c["z_c"], c["gamma"],
                      stretching_type=stretching, order=order)
    sl.set_tanh_g(plan.tanh_g)
    mids = [cuda(a) for a in c["mids"]]
    dt = c["dt"]
    for ic, comp in enumerate("uvw"):
        n = plan.n_arrival(comp)
        xd, yd, zd = (torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(3))
        plan.departure(comp, *mids, dt, xd, yd, zd)
        ref = sl.departure(comp, *c["mids"], dt)
        ok, same, dev = feet_ok([t.cpu().numpy() for t in (xd, yd, zd)], ref, c["Lx"])
        check(f"feet {comp} vs oracle", ok,
              f"identical={int(same.sum())}/{same.size} max|d|={dev:.1e}")
        ix, iy, k0 = (torch.empty(n, dtype=torch.int32, device="cuda") for _ in range(3))
        gx, gy, gz = (cuda(a) for a in ref)
        plan.stencil(comp, gx, gy, gz, ix, iy, k0)
        vals, _, (rix, riy, rk0) = sl.gather(comp, c["fields"][ic], *ref, want_index=True)
        same_idx = ((ix.cpu().numpy() == rix) & (iy.cpu().numpy() == riy)
                    & (k0.cpu().numpy() == rk0))
        check(f"stencil {comp} vs oracle", same_idx.all(),
              f"identical={int(same_idx.sum())}/{same_idx.size}")
        out = torch.empty(n, dtype=torch.float64, device="cuda")
        plan.gather(comp, cuda(c["fields"][ic]), gx, gy, gz, out)
        e = rel_err(out, vals)
        check(f"gather {comp} vs oracle", e < 1e-13, f"rel={e:.1e}")
    shapes = plan.shapes
    outs = [torch.zeros(shapes[k], dtype=torch.float64, device="cuda") for k in "uvw"]
    ncl = torch.zeros(1, dtype=torch.int64, device="cuda")
    plan.advect(*[cuda(a) for a in c["fields"]], *mids, dt, *outs, ncl)
    ref_out = sl.advect(*c["fields"], *c["mids"], dt)
    for nm, a, b in zip("uvw", outs, ref_out):
        e = rel_err(a, b)
        check(f"advect {nm} vs oracle", e < 1e-13, f"rel={e:.1e}")
    check("clamp count vs oracle", int(ncl.item()) == sl.n_clamped_last,
          f"{int(ncl.item())} vs {sl.n_clamped_last}")


@pytest.mark.parametrize("tag", ["sym", "bot"])
def test_solves_vs_reference(check, golden, tag):
    g = golden("solve_small")
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    gg = dict(nx=nx, ny=ny, nz=nz, Lx=g["Lx"], Ly=g["Ly"], Lz=g["Lz"], gamma=g["gamma"],
              z_f=g[f"{tag}_z_f"], z_c=g[f"{tag}_z_c"],
              stretching="symmetric" if tag == "sym" else "bottom")
    top = str(g[f"{tag}_top"])
    plan = make_plan(gg, order=4, nu=float(g[f"{tag}_nu"]), top=top)
    dt = float(g[f"{tag}_dt"])
    for theta, th in ((1.0, "t1"), (0.5, "t05")):
        for comp in "uvw":
            f = cuda(g[f"{tag}_{comp}"])
            plan.zsolve(comp, f, dt, theta)
            e = rel_err(f, g[f"{tag}_{th}_{comp}"])
            check(f"zsolve {comp} theta={theta}", e < 1e-13, f"rel={e:.1e}")
    p = torch.zeros(plan.shapes["p"], dtype=torch.float64, device="cuda")
    plan.poisson(cuda(g[f"{tag}_div"]), p, 1.0)
    e = rel_err(p, g[f"{tag}_p"])
    check("poisson p vs reference", e < 1e-12, f"rel={e:.1e}")


@pytest.mark.parametrize("nz", [60, 120, 250, 500, 1000])
def test_zsolve_column_lengths(check, nz):
    """z-solves whose columns hold 1, 2, 4, 8 and 16 unknowns per lane (the
    register-scan kernel's variants) against the oracle's PCR solve
    (operators.py:382-630, tridiag.py:36-61), theta = 1 and 1/2, both top-wall
    conditions: within 1e-12 of max|u| (Thomas against PCR rounding)."""
    from slchannel_amd.grid import generate_grid

    nx, ny, gamma, Lz, nu, dt = 3, 5, 1.8, 2.0, 1.0 / 2000.0, 0.05
    z_f, z_c, dz_f, dz_c = (t.numpy() for t in generate_grid(gamma, nz, Lz))
    rng = np.random.default_rng(nz)
    for top in ("dirichlet", "neumann"):
        gg = dict(nx=nx, ny=ny, nz=nz, Lx=4 * math.pi, Ly=2 * math.pi, Lz=Lz, gamma=gamma,
                  z_f=z_f, z_c=z_c, stretching="symmetric")
        plan = make_plan(gg, order=4, nu=nu, top=top)
        for theta in (1.0, 0.5):
            for comp in "uvw":
                a = rng.normal(size=plan.shapes[comp])
                f = cuda(a)
                plan.zsolve(comp, f, dt, theta)
                if comp == "w":
                    ref = orc.solve_implicit_diffusion_w(a, dt, nx, ny, nz, dz_c, dz_f, nu, theta)
                else:
                    ref = orc.solve_implicit_diffusion_u(a, dt, nx, ny, nz, dz_c, dz_f, nu, theta,
                                                         top)
                e = rel_err(f, ref)
                check(f"nz={nz} {top} theta={theta} {comp}", e < 1e-12, f"rel={e:.1e}")


@pytest.mark.parametrize("nz", [40, 300, 600, 1100, 2100])
def test_poisson_column_lengths(check, nz):
    """FFT-Poisson solve (projection.py:24-247) at column lengths that take
    the partition solve with 32, 64, 128 and 256 segments per mode (256- and
    512-thread workgroups) and, past 2048 levels, the sequential sweep:
    against the oracle's solve, both top walls.  Bar: 64 nz eps of max|p|,
    the forward-error form of an elimination over nz levels (the (0,0)
    mode's pinned system dominates); the sequential sweep measured 5e-13
    (nz 256), 1.6e-12 (600) and 2.6e-12 (1100) on these cases, the partition
    solve 4e-13, 3.0e-12 and 1.6e-12."""
    nx, ny, Lz = 6, 5, 2.0
    Lx, Ly = 4 * math.pi, 2 * math.pi
    dx, dy = Lx / nx, Ly / ny
    z_f, z_c, dz_f, dz_c = orc.generate_grid(1.5, nz, Lz)
    rng = np.random.default_rng(nz)
    for top in ("dirichlet", "neumann"):
        gg = dict(nx=nx, ny=ny, nz=nz, Lx=Lx, Ly=Ly, Lz=Lz, gamma=1.5, z_f=z_f, z_c=z_c,
                  stretching="symmetric")
        plan = make_plan(gg, order=4, top=top)
        div = rng.normal(size=(nx, ny, nz))
        div -= div.mean()
        p = torch.zeros(plan.shapes["p"], dtype=torch.float64, device="cuda")
        plan.poisson(cuda(div), p, 1.0)
        fft = orc.initialize_fft_solver(nx, ny, nz, dx, dy, dz_c, dz_f, top)
        ref = orc.solve_poisson_fft(div, fft)
        e = rel_err(p, ref)
        tol = 64 * nz * EPS
        check(f"nz={nz} {top}: p vs oracle within {tol:.1e}", e < tol, f"rel={e:.1e}")


def test_divergence_projection_bulk_vs_oracle(check):
    c = _random_case(24, 20, 28, 4, "symmetric", seed=9)
    plan = make_plan(c)
    nx, ny, nz = c["nx"], c["ny"], c["nz"]
    dx, dy = c["Lx"] / nx, c["Ly"] / ny
    z_f, z_c = c["z_f"], c["z_c"]
    dz_f, dz_c = z_f[1:] - z_f[:-1], z_c[1:] - z_c[:-1]
    u, v, w = (a.copy() for a in c["fields"])
    tu, tv, tw = cuda(u), cuda(v), cuda(w)
    div = torch.empty(nx, ny, nz, dtype=torch.float64, device="cuda")
    plan.divergence(tu, tv, tw, div)
    ref = orc.compute_divergence(u, v, w, nx, ny, nz, dx, dy, dz_f)
    check("divergence bit-exact", np.array_equal(div.cpu().numpy(), ref), "")
    rng = np.random.default_rng(3)
    p = rng.normal(size=(nx + 2, ny + 2, nz + 2))
    plan.project(tu, tv, tw, cuda(p), 0.37)
    orc.project_velocity(u, v, w, p, nx, ny, nz, dx, dy, dz_c, dz_f, 0.37)
    for nm, a, b in (("u", tu, u), ("v", tv, v), ("w", tw, w)):
        check(f"project {nm} bit-exact", np.array_equal(a.cpu().numpy(), b), "")
    out = torch.zeros(4, dtype=torch.float64, device="cuda")
    plan.bulk_velocity(tu, out)
    cv = np.broadcast_to((dx * dy * dz_f).reshape(1, 1, -1), (nx, ny, nz))
    ub = orc.compute_bulk_velocity(u, cv, c["Lx"] * c["Ly"] * c["Lz"])
    e = abs(out[3].item() - ub) / abs(ub)
    check("bulk velocity", e < 1e-14, f"rel={e:.1e}")
    # fused projection step: div -> rocFFT -> Thomas -> iFFT -> grad -> ghosts
    u2, v2, w2 = (a.copy() for a in c["fields"])
    t2 = [cuda(a) for a in (u2, v2, w2)]
    pp = torch.zeros(plan.shapes["p"], dtype=torch.float64, device="cuda")
    plan.project_step(*t2, 0.5, 0.5, pp)
    fft = orc.initialize_fft_solver(nx, ny, nz, dx, dy, dz_c, dz_f)
    d2 = orc.compute_divergence(u2, v2, w2, nx, ny, nz, dx, dy, dz_f)
    phi = orc.solve_poisson_fft(d2 / 0.5, fft)
    orc.project_velocity(u2, v2, w2, phi, nx, ny, nz, dx, dy, dz_c, dz_f, 0.5)
    orc.apply_bc_all(u2, v2, w2)
    for nm, a, b in (("u", t2[0], u2), ("v", t2[1], v2), ("w", t2[2], w2), ("p", pp, phi)):
        e = rel_err(a, b)
        check(f"project_step {nm}", e < 1e-12, f"rel={e:.1e}")
    d3 = torch.empty(nx, ny, nz, dtype=torch.float64, device="cuda")
    plan.divergence(*t2, d3)
    md = d3.abs().max().item()
    check("post-projection max|div|", md < 1e-11, f"{md:.2e}")


def _flow_from_golden(g):
    from slchannel_amd import SLChannelFlow

    cfg = {
        "grid": {"nx": int(g["nx"]), "ny": int(g["ny"]), "nz": int(g["nz"])},
        "domain": {"Lx": float(g["Lx"]), "Ly": float(g["Ly"]), "Lz": float(g["Lz"]),
                   "stretching_type": str(g["stretching"])},
        "flow": {"Re": float(g["Re"]), "U_bulk": 1.0, "gamma": float(g["gamma"])},
        "boundary_conditions": {"top_wall": {"type": str(g["top"])}},
        "time": {"dt": float(g["dt"])},
        "sl": {"interp_order": int(g["order"]), "n_traj_iters": 2,
               "bdf2_pressure": str(g["bdf2_pressure"]) if "bdf2_pressure" in g else "noninc",
               "bdf2_xy_rhs": str(g["bdf2_xy_rhs"]) if "bdf2_xy_rhs" in g else "extrap",
               "traj_interp_order": int(g["traj_order"]) if "traj_order" in g else 2},
    }
    return SLChannelFlow(config=cfg, z_tables=(g["z_f"], g["z_c"]))


@pytest.mark.parametrize("name", ["o6", "o4_bottom_neumann", "o4_inc", "o6_lagged", "o4_traj4"])
def test_step_vs_reference(check, golden, name):
    g = golden("step_" + name)
    flow = _flow_from_golden(g)
    flow.set_fields(g["u0"], g["v0"], g["w0"], g["p0"])
    dt = float(g["dt"])
    for n in range(int(g["n_steps"])):
        ub, fo = flow.step_sl_bdf2(dt)
        check(f"step {n + 1} u_bulk", abs(ub.item() - g["u_bulk"][n]) < 1e-14,
              f"{ub.item()!r}")
        check(f"step {n + 1} forcing",
              abs(fo.item() - g["forcing"][n]) < 1e-12 * max(1.0, abs(g["forcing"][n])),
              f"{fo.item():.6e} vs {g['forcing'][n]:.6e}")
        check(f"step {n + 1} n_clamped", int(flow.sl.n_clamped_last.item()) == g["n_clamped"][n],
              "")
        if f"u_{n + 1}" in g:
            for nm, a in (("u", flow.u), ("v", flow.v), ("w", flow.w), ("p", flow.p)):
                e = rel_err(a, g[f"{nm}_{n + 1}"])
                check(f"step {n + 1} {nm}", e < 1e-12, f"rel={e:.1e}")


def _config(n, order=6, Re=2800.0, nz=None):
    return {
        "grid": {"nx": n, "ny": n, "nz": nz or n},
        "domain": {"Lx": 4 * math.pi, "Ly": 2 * math.pi, "Lz": 2.0},
        "flow": {"Re": Re, "U_bulk": 1.0, "gamma": 1.5},
        "sl": {"interp_order": order},
    }


@pytest.mark.parametrize("order", [4, 6])
def test_step_vs_oracle_seeded(check, order):
    """Seeded solenoidal init, BDF1 + BDF2 steps, HIP vs oracle every step."""
    from slchannel_amd import SLChannelFlow

    flow = SLChannelFlow(config=_config(32, order=order, nz=24))
    flow.initialize(seed=1001, urms=0.1, kmax=6.0)
    co = orc.ChannelOracle(flow.nx, flow.ny, flow.nz, flow.Lx, flow.Ly, flow.Lz, flow.gamma,
                           2800.0, order=order, z_f=flow.z_f.numpy(), z_c=flow.z_c.numpy(),
                           tanh_g=flow.plan.tanh_g)
    co.set_fields(flow.u.cpu().numpy(), flow.v.cpu().numpy(), flow.w.cpu().numpy(),
                  flow.p.cpu().numpy())
    from slchannel_amd.diagnostics import cfl_dt

    dt = cfl_dt(flow, 3.0)  # trajectory CFL 3 (operators.py:839-940)
    worst, per_step = 0.0, []
    for n in range(8):
        ub, fo = flow.step_sl_bdf2(dt)
        rub, rfo = co.step_sl_bdf2(dt)
        e = max(rel_err(a, b) for a, b in ((flow.u, co.u), (flow.v, co.v), (flow.w, co.w)))
        per_step.append(f"{e:.0e}")
        worst = max(worst, e)
        md = flow.divergence().abs().max().item()
        check(f"step {n + 1} max|div|", md < 1e-11, f"{md:.2e}")
        check(f"step {n + 1} u_bulk", abs(ub.item() - 1.0) < 1e-14, f"{ub.item()!r}")
        check(f"step {n + 1} clamps vs oracle",
              int(flow.sl.n_clamped_last.item()) == co.n_clamped_last,
              f"{int(flow.sl.n_clamped_last.item())} vs {co.n_clamped_last}")
    check("fields vs oracle over 8 steps", worst < 1e-12,
          f"max rel={worst:.1e} per step={per_step} dt={dt:.4f}")


@pytest.mark.parametrize("order", [4, 6])
def test_tiled_gather_matches_direct(check, order):
    """The LDS-tiled gather must reproduce the direct gather bit for bit
    (same stencils, same accumulation order), for every tile shape."""
    c = _random_case(40, 36, 48, order, "symmetric", seed=31, cfl=3.0)
    plan = make_plan(c)
    mids = [cuda(a) for a in c["mids"]]
    for ic, comp in enumerate("uvw"):
        n = plan.n_arrival(comp)
        xd, yd, zd = (torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(3))
        plan.departure(comp, *mids, 2.0 * c["dt"], xd, yd, zd)
        f = cuda(c["fields"][ic])
        ref = torch.empty(n, dtype=torch.float64, device="cuda")
        ncl0 = torch.zeros(1, dtype=torch.int64, device="cuda")
        plan.set_option(plan.OPT_GATHER_TILE, -1)
        plan.gather(comp, f, xd, yd, zd, ref, ncl0)
        for shape in range(4):
            plan.set_option(plan.OPT_GATHER_TILE, shape)
            out = torch.empty(n, dtype=torch.float64, device="cuda")
            ncl = torch.zeros(1, dtype=torch.int64, device="cuda")
            plan.gather(comp, f, xd, yd, zd, out, ncl)
            check(f"{comp} shape {shape} bit-identical", torch.equal(out, ref) and
                  int(ncl.item()) == int(ncl0.item()),
                  f"max|d|={(out - ref).abs().max().item():.1e} ncl={int(ncl.item())}")
        # the field 8 bytes past a 16-byte boundary: boxes staged through
        # registers instead of the direct-to-LDS pairs, same doubles
        buf = torch.empty(f.numel() + 1, dtype=torch.float64, device="cuda")
        fo = buf[1:].view(f.shape)
        fo.copy_(f)
        plan.set_option(plan.OPT_GATHER_TILE, 0)
        out = torch.empty(n, dtype=torch.float64, device="cuda")
        plan.gather(comp, fo, xd, yd, zd, out)
        check(f"{comp} field at an 8-byte offset: bit-identical",
              fo.data_ptr() % 16 == 8 and torch.equal(out, ref),
              f"ptr%16={fo.data_ptr() % 16} max|d|={(out - ref).abs().max().item():.1e}")


@pytest.mark.parametrize("order", [4, 6])
def test_gather_wide_boxes(check, order):
    """Feet displaced at random by up to +-5 cells in x and y (the large-box
    workload's isolated gather): the tiles' stencil boxes exceed the default
    LDS budget, so the first pass defers them and a second pass with a
    larger budget gathers them.  Bit-identical to the direct gather, within
    1e-13 of the oracle, every arrival point written once."""
    nx, ny, nz = 48, 40, 36
    c = _random_case(nx, ny, nz, order, "symmetric", seed=53, cfl=1.0)
    plan = make_plan(c)
    dx, dy = c["Lx"] / nx, c["Ly"] / ny
    sl = orc.SLOracle(nx, ny, nz, dx, dy, c["Lx"], c["Ly"], c["Lz"], c["z_f"], c["z_c"],
                      c["gamma"], order=order)
    sl.set_tanh_g(plan.tanh_g)
    rng = np.random.default_rng(54)
    for ic, comp in enumerate("uvw"):
        n = plan.n_arrival(comp)
        nzi = nz - 1 if comp == "w" else nz
        i = np.arange(n) // (ny * nzi) + 1.0
        j = (np.arange(n) // nzi) % ny + 1.0
        k = np.arange(n) % nzi
        xd = (i - (0.0 if comp == "u" else 0.5) + rng.uniform(-5, 5, n)) * dx
        yd = (j - (0.5 if comp != "v" else 0.0) + rng.uniform(-5, 5, n)) * dy
        zd = (c["z_f"][1:nz] if comp == "w" else c["z_c"][1:nz + 1])[k]
        f = cuda(c["fields"][ic])
        dx_, dy_, dz_ = (cuda(a) for a in (xd, yd, zd))
        plan.set_option(plan.OPT_GATHER_TILE, -1)
        ref = torch.empty(n, dtype=torch.float64, device="cuda")
        plan.gather(comp, f, dx_, dy_, dz_, ref)
        plan.set_option(plan.OPT_GATHER_TILE, 0)
        out = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
        plan.gather(comp, f, dx_, dy_, dz_, out)
        check(f"o{order} {comp}: two-pass gather bit-identical to the direct gather",
              torch.equal(out, ref), f"max|d|={(out - ref).abs().max().item():.1e}")
        o, _ = sl.gather(comp, c["fields"][ic], xd, yd, zd)
        e = rel_err(out, o)
        check(f"o{order} {comp}: within 1e-13 of the oracle", e < 1e-13, f"{e:.1e}")


@pytest.mark.parametrize("order", [4, 6])
def test_gather_wide_mixed_tiles(check, order):
    """Feet displaced by up to +-5 cells on a checkerboard of 4 x 4 columns
    (the first pass's tiles) and by under a cell elsewhere, 3% of them
    outside the channel in z: the second pass's wider tiles overlap deferred
    and gathered first-pass tiles alike.  Output and clamp count identical to
    the direct gather (every point gathered and counted once), within 1e-13 of
    the oracle."""
    nx, ny, nz = 40, 44, 30
    c = _random_case(nx, ny, nz, order, "symmetric", seed=57, cfl=1.0)
    plan = make_plan(c)
    dx, dy = c["Lx"] / nx, c["Ly"] / ny
    sl = orc.SLOracle(nx, ny, nz, dx, dy, c["Lx"], c["Ly"], c["Lz"], c["z_f"], c["z_c"],
                      c["gamma"], order=order)
    sl.set_tanh_g(plan.tanh_g)
    rng = np.random.default_rng(58)
    for ic, comp in enumerate("uvw"):
        n = plan.n_arrival(comp)
        nzi = nz - 1 if comp == "w" else nz
        ii = np.arange(n) // (ny * nzi)
        jj = (np.arange(n) // nzi) % ny
        k = np.arange(n) % nzi
        amp = np.where((ii // 4 + jj // 4) % 2 == 0, 5.0, 0.4)
        xd = (ii + 1.0 - (0.0 if comp == "u" else 0.5) + amp * rng.uniform(-1, 1, n)) * dx
        yd = (jj + 1.0 - (0.5 if comp != "v" else 0.0) + amp * rng.uniform(-1, 1, n)) * dy
        zd = (c["z_f"][1:nz] if comp == "w" else c["z_c"][1:nz + 1])[k].copy()
        out_z = rng.random(n) < 0.03
        zd[out_z] = np.where(rng.random(out_z.sum()) < 0.5, -0.05, c["Lz"] + 0.05)
        f = cuda(c["fields"][ic])
        dx_, dy_, dz_ = (cuda(a) for a in (xd, yd, zd))
        plan.set_option(plan.OPT_GATHER_TILE, -1)
        ref = torch.empty(n, dtype=torch.float64, device="cuda")
        ncl0 = torch.zeros(1, dtype=torch.int64, device="cuda")
        plan.gather(comp, f, dx_, dy_, dz_, ref, ncl0)
        plan.set_option(plan.OPT_GATHER_TILE, 0)
        out = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
        ncl = torch.zeros(1, dtype=torch.int64, device="cuda")
        plan.gather(comp, f, dx_, dy_, dz_, out, ncl)
        check(f"o{order} {comp}: two-pass gather bit-identical to the direct gather",
              torch.equal(out, ref), f"max|d|={(out - ref).abs().max().item():.1e}")
        o, ocl = sl.gather(comp, c["fields"][ic], xd, yd, zd)
        check(f"o{order} {comp}: clamp count {int(ncl.item())} == direct "
              f"{int(ncl0.item())} == oracle {ocl} == clamped feet {int(out_z.sum())}",
              int(ncl.item()) == int(ncl0.item()) == ocl == int(out_z.sum()))
        e = rel_err(out, o)
        check(f"o{order} {comp}: within 1e-13 of the oracle", e < 1e-13, f"{e:.1e}")


@pytest.mark.parametrize("order,shape,cfl",[(4, (40, 36, 48), 3.0), (6, (40, 36, 48), 3.0),
                                             (6, (37, 29, 23), 3.0), (6, (40, 36, 48), 12.0)],
                         ids=["o4", "o6", "o6-ragged", "o6-cfl12"])
def test_tiled_step_kernels_match_direct(check, order, shape, cfl):
    """advect, BDF1 and BDF2 predictors: LDS-tiled kernels (default, with the
    fused one-pass BDF2 predictor) against the direct global-load kernels and
    against the split far/near BDF2 passes -- identical results and clamp
    counts.  The ragged grid leaves partial tiles in x, y and z; at CFL 12
    the stencil boxes outgrow the LDS budget and the tiles take the
    direct-load branch of the tile kernels."""
    c = _random_case(*shape, order, "symmetric", seed=17, cfl=cfl)
    plan = make_plan(c, nu=1e-3)
    f = [cuda(a) for a in c["fields"]]
    m = [cuda(a) for a in c["mids"]]
    dt = c["dt"]
    res = {}
    for shape in (-1, 0, "split"):
        plan.set_option(plan.OPT_GATHER_TILE, 0 if shape == "split" else shape)
        plan.set_option(plan.OPT_BDF2_FUSED, 0 if shape == "split" else 1)
        outs = {}
        for name in ("advect", "bdf1", "bdf2"):
            o = [torch.zeros(plan.shapes[k], dtype=torch.float64, device="cuda") for k in "uvw"]
            ncl = torch.zeros(1, dtype=torch.int64, device="cuda")
            if name == "advect":
                plan.advect(*f, *m, dt, *o, ncl)
            elif name == "bdf1":
                plan.bdf1_predict(*f, dt, *o, ncl)
            else:
                plan.bdf2_predict(*f, *m, dt, *o, ncl)
            outs[name] = (o, int(ncl.item()))
        res[shape] = outs
    plan.set_option(plan.OPT_GATHER_TILE, 0)
    plan.set_option(plan.OPT_BDF2_FUSED, 1)
    for name in ("advect", "bdf1", "bdf2"):
        (a, na), (b, nb) = res[0][name], res[-1][name]
        same = all(torch.equal(x, y) for x, y in zip(a, b))
        dev = max((x - y).abs().max().item() for x, y in zip(a, b))
        check(f"{name}: tiled == direct", same and na == nb, f"max|d|={dev:.1e} ncl={na},{nb}")
    (a, na), (b, nb) = res[0]["bdf2"], res["split"]["bdf2"]
    same = all(torch.equal(x, y) for x, y in zip(a, b))
    dev = max((x - y).abs().max().item() for x, y in zip(a, b))
    check("bdf2: fused == split", same and na == nb, f"max|d|={dev:.1e} ncl={na},{nb}")


@pytest.mark.parametrize("stretching,gamma,nz", [("symmetric", 1.5, 256), ("symmetric", 2.6, 180),
                                                 ("bottom", 1.8, 512)])
def test_fast_z_locate_equals_analytic(check, stretching, gamma, nz):
    """The kernels' table-refined z-locate must equal the analytic inverse-map
    path everywhere, including points within ulps of every face; the analytic
    path must equal the oracle's off exact ties."""
    from slchannel_amd import HipPlan
    from slchannel_amd.grid import generate_grid

    z_f, z_c, _, _ = (t.numpy() for t in generate_grid(gamma, nz, 2.0, stretching_type=stretching))
    plan = HipPlan(8, 8, nz, 4.0, 4.0, 2.0, z_f, z_c, gamma, stretching_type=stretching, order=6)
    dz = np.diff(z_f)
    pts = [np.random.default_rng(0).uniform(0.0, 2.0, 200000), z_c, z_f]
    for k_ulp in (1, 2, 4, 16, 1000):
        pts += [np.nextafter(z_f, 10.0 * np.ones_like(z_f)) if k_ulp == 1 else
                z_f + k_ulp * np.spacing(z_f), z_f - k_ulp * np.spacing(z_f)]
    for rel in (1e-13, 1e-12, 1e-11, 1e-10, 3e-10, 1e-9):
        pts += [z_f[:-1] + rel * dz, z_f[1:] - rel * dz]
    z = np.concatenate(pts)
    tz = cuda(z)
    for comp in ("u", "w"):
        ex = torch.empty(z.size, dtype=torch.int32, device="cuda")
        fa = torch.empty_like(ex)
        plan.locate_z(comp, tz, ex, fa)
        ex, fa = ex.cpu().numpy(), fa.cpu().numpy()
        check(f"{comp}: fast == analytic on {z.size} points", np.array_equal(ex, fa),
              f"mismatches={int((ex != fa).sum())}")
        sl = orc.SLOracle(8, 8, nz, 0.5, 0.5, 4.0, 4.0, 2.0, z_f, z_c, gamma,
                          stretching_type=stretching, order=6)
        sl.set_tanh_g(plan.tanh_g)
        ref = sl.locate_z(comp, z)
        kf = sl.face_coord(np.clip(z, sl.z_lo, sl.z_hi))
        tie = np.abs(kf - np.round(kf)) < 1e-9
        off = ~tie
        check(f"{comp}: analytic == oracle off ties", np.array_equal(ex[off], ref[off]),
              f"mismatches={int((ex[off] != ref[off]).sum())} ties={int(tie.sum())}")


def _lebesgue(order, n=2001):
    """1-D Lebesgue constant of the order-point Lagrange stencil on uniform
    nodes over its central interval (where the gathers evaluate it)."""
    nodes = np.arange(order) - (order // 2 - 1)
    t = np.linspace(0.0, 1.0, n)
    tot = np.zeros_like(t)
    for m in range(order):
        lm = np.ones_like(t)
        for l in range(order):
            if l != m:
                lm *= (t - nodes[l]) / (nodes[m] - nodes[l])
        tot += np.abs(lm)
    return float(tot.max())


@pytest.mark.parametrize("order", [4, 6])
def test_fp32_storage_error_per_field(check, order):
    """FP32-storage gathers (float fields, FP64 weights/accumulation/feet)
    against the FP64 path: same clamp counts (the feet do not change), and
    per field |d| <= 2^-24 * Lambda^3 * max|F| (times 4/3 + 1/3 for the BDF2
    predictor), Lambda the 1-D Lebesgue constant of the stencil (x1.1 for the
    stretched z nodes).  The error must also be non-zero: the float copies
    are really read."""
    c = _random_case(40, 36, 48, order, "symmetric", seed=71, cfl=3.0)
    plan = make_plan(c, nu=1e-3)
    f = [cuda(a) for a in c["fields"]]
    m = [cuda(a) for a in c["mids"]]
    dt = c["dt"]
    lam = _lebesgue(order)
    base = 2.0**-24 * lam**3 * 1.1
    res = {}
    for bits in (64, 32):
        plan.set_option(plan.OPT_FIELD_STORAGE, bits)
        outs = {}
        for name in ("advect", "bdf1", "bdf2"):
            o = [torch.zeros(plan.shapes[k], dtype=torch.float64, device="cuda") for k in "uvw"]
            ncl = torch.zeros(1, dtype=torch.int64, device="cuda")
            if name == "advect":
                plan.advect(*f, *m, dt, *o, ncl)
            elif name == "bdf1":
                plan.bdf1_predict(*f, dt, *o, ncl)
            else:
                plan.bdf2_predict(*f, *m, dt, *o, ncl)
            outs[name] = (o, int(ncl.item()))
        res[bits] = outs
    # float storage, split far/near BDF2 passes: the fused pass's doubles
    plan.set_option(plan.OPT_BDF2_FUSED, 0)
    o = [torch.zeros(plan.shapes[k], dtype=torch.float64, device="cuda") for k in "uvw"]
    ncl = torch.zeros(1, dtype=torch.int64, device="cuda")
    plan.bdf2_predict(*f, *m, dt, *o, ncl)
    plan.set_option(plan.OPT_BDF2_FUSED, 1)
    plan.set_option(plan.OPT_FIELD_STORAGE, 64)
    a32, n32 = res[32]["bdf2"]
    check("fp32 bdf2: fused == split",
          all(torch.equal(x, y) for x, y in zip(a32, o)) and n32 == int(ncl.item()),
          f"max|d|={max((x - y).abs().max().item() for x, y in zip(a32, o)):.1e}")
    for name in ("advect", "bdf1", "bdf2"):
        (a, na), (b, nb) = res[32][name], res[64][name]
        scale = base * ((4.0 / 3.0 + 1.0 / 3.0) if name == "bdf2" else 1.0)
        srcs = c["fields"] if name != "bdf2" else [np.maximum(np.abs(x), np.abs(y)) for x, y in
                                                    zip(c["fields"], c["mids"])]
        for comp, x, y, s in zip("uvw", a, b, srcs):
            d = (x - y).abs().max().item()
            bound = scale * float(np.abs(s).max())
            check(f"{name} {comp}: fp32 storage vs fp64", 0.0 < d <= bound,
                  f"max|d|={d:.2e} bound={bound:.2e} rel={d / float(np.abs(s).max()):.1e}")
        check(f"{name}: clamp counts unchanged", na == nb, f"{na} vs {nb}")


def test_fp32_storage_step_vs_oracle(check):
    """interp_dtype fp32_accum64 (FP32-storage gathers) over 8 BDF1/BDF2
    steps of the seeded order-6 case against the FP64 oracle: per-field
    error stated, below 1e-6 relative; the FP64 projection still gives
    max|div| < 1e-11 and u_bulk = 1.0 every step."""
    from slchannel_amd import SLChannelFlow
    from slchannel_amd.diagnostics import cfl_dt

    cfg = _config(32, order=6, nz=24)
    cfg["sl"]["interp_dtype"] = "fp32_accum64"
    flow = SLChannelFlow(config=cfg)
    flow.initialize(seed=1001, urms=0.1, kmax=6.0)
    co = orc.ChannelOracle(flow.nx, flow.ny, flow.nz, flow.Lx, flow.Ly, flow.Lz, flow.gamma,
                           2800.0, order=6, z_f=flow.z_f.numpy(), z_c=flow.z_c.numpy(),
                           tanh_g=flow.plan.tanh_g)
    co.set_fields(flow.u.cpu().numpy(), flow.v.cpu().numpy(), flow.w.cpu().numpy(),
                  flow.p.cpu().numpy())
    dt = cfl_dt(flow, 3.0)
    errs = {k: 0.0 for k in "uvwp"}
    for n in range(8):
        ub, _ = flow.step_sl_bdf2(dt)
        co.step_sl_bdf2(dt)
        for k, a, b in (("u", flow.u, co.u), ("v", flow.v, co.v), ("w", flow.w, co.w),
                        ("p", flow.p, co.p)):
            errs[k] = max(errs[k], rel_err(a, b))
        md = flow.divergence().abs().max().item()
        check(f"step {n + 1} max|div|", md < 1e-11, f"{md:.2e}")
        check(f"step {n + 1} u_bulk", abs(ub.item() - 1.0) < 1e-14, f"{ub.item()!r}")
    for k in "uvw":
        check(f"{k}: fp32 storage vs fp64 oracle over 8 steps", 0.0 < errs[k] < 1e-6,
              f"max rel={errs[k]:.2e}")
    check("p: fp32 storage vs fp64 oracle over 8 steps", errs["p"] < 1e-5,
          f"max rel={errs['p']:.2e}")
