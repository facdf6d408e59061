"""Write the golden input/output vectors that pin the oracle.

Run ONCE in the build container (the reference checkout is not present on the
GPU machines):

    MPLBACKEND=Agg python tests/golden/make_golden.py /root/reference [name ...]

(names, e.g. ``step_o4_inc``, regenerate only those files.)

It imports the reference package (gmcavallazzi/SLchannel, pure Python/torch)
from the given checkout and evaluates, on small seeded inputs, exactly the
functions whose results no closed form or invariant fixes bit-for-bit:

* ``grid_*.npz``     generate_grid (utils.py:21-61), the SLAdvector tables
                     (semilag.py:102-127, 240-247) and _face_coord
                     (semilag.py:400-411) at faces and random points;
* ``advect_*.npz``   SLAdvector.advect (semilag.py:584-671) on the eager fp64
                     path, with the departure feet (semilag.py:534-552) and the
                     stencil base indices (semilag.py:471-500) per component;
* ``solve_*.npz``    solve_poisson_fft (projection.py:133-197) and
                     solve_implicit_diffusion_{u,v,w} (operators.py:382-630);
* ``step_*.npz``     SLChannelFlow.step_sl_bdf2 (solver.py:813-1018), BDF1
                     bootstrap included, from a seeded projected state; two
                     cases with the non-default BDF2 options (bdf2_pressure
                     'inc', bdf2_xy_rhs 'lagged', solver.py:256-281);
* ``*_traj4.npz``    advect and step with tricubic trajectory sampling
                     (traj_interp_order 4, semilag.py:179-209, 512-528);
* ``imex_small.npz`` compute_momentum_rhs_fused_imex (operators.py:639-835) on
                     both stretchings, and three SLChannelFlow.step_imex steps
                     (solver.py:1055-1134) with that right-hand side (the
                     path the reference takes on a GPU, solver.py:1027-1039);
* ``stats_small.npz`` TurbulenceStats.accumulate_statistics over three
                     snapshots, with two extra spectra planes, and
                     finalize_statistics (turbstats.py:35-467);
* ``diag_small.npz`` compute_u_tau (utils.py:116-146, both top-wall types)
                     and compute_cfl_fused (operators.py:839-940).

Inputs are generated here with numpy (seeded) and stored next to the outputs,
so the fixtures are self-contained: data only, no reference code.
"""

import math
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _import_reference(path):
    sys.path.insert(0, path)
    import torch

    import slchannel  # noqa: F401  (sets float64 default)
    from slchannel import operators, projection, semilag, solver, utils

    return torch, semilag, solver, utils, projection, operators


def smooth_random(rng, X, Y, Z, Lx, Ly, Lz, n_modes=4, amp=1.0):
    """Sum of a few random smooth modes evaluated at broadcast positions."""
    out = 0.0 * (X + Y + Z)
    for _ in range(n_modes):
        mx = rng.integers(0, 3)
        my = rng.integers(0, 3)
        lz = rng.uniform(0.3, 2.0)
        a = rng.normal() * amp / n_modes
        px, py, pz = rng.uniform(0, 2 * math.pi, 3)
        out = out + a * np.sin(2 * math.pi * mx * X / Lx + px) * np.cos(
            2 * math.pi * my * Y / Ly + py
        ) * np.cos(math.pi * lz * Z / Lz + pz)
    return out


def positions(comp, nx, ny, dx, dy, z_f, z_c):
    """Physical coordinates of every entry of the ghost-shaped array."""
    if comp == "u":
        X = (np.arange(nx + 1) * dx)[:, None, None]
        Y = ((np.arange(ny + 2) - 0.5) * dy)[None, :, None]
        Z = z_c[None, None, :]
    elif comp == "v":
        X = ((np.arange(nx + 2) - 0.5) * dx)[:, None, None]
        Y = (np.arange(ny + 1) * dy)[None, :, None]
        Z = z_c[None, None, :]
    else:
        X = ((np.arange(nx + 2) - 0.5) * dx)[:, None, None]
        Y = ((np.arange(ny + 2) - 0.5) * dy)[None, :, None]
        Z = z_f[None, None, :]
    return X, Y, Z


def make_fields(rng, nx, ny, nz, dx, dy, Lx, Ly, Lz, z_f, z_c, base=None, amp=1.0, wamp=None):
    out = []
    for comp in "uvw":
        X, Y, Z = positions(comp, nx, ny, dx, dy, z_f, z_c)
        a = amp if (comp != "w" or wamp is None) else wamp
        f = smooth_random(rng, X, Y, Z, Lx, Ly, Lz, amp=a)
        if base is not None and comp in base:
            f = f + base[comp](X, Y, Z)
        shape = {"u": (nx + 1, ny + 2, nz + 2), "v": (nx + 2, ny + 1, nz + 2),
                 "w": (nx + 2, ny + 2, nz + 1)}[comp]
        out.append(np.ascontiguousarray(np.broadcast_to(f, shape), dtype=np.float64))
    return out


def gen_grid(ref, name, gamma, nz, Lz, stretching):
    torch, semilag, solver, utils, projection, operators = ref
    z_f, z_c, dz_f, dz_c = utils.generate_grid(gamma, nz, Lz, stretching_type=stretching)
    adv = semilag.SLAdvector(8, 6, nz, 0.5, 0.5, 4.0, 3.0, Lz, z_f, z_c, gamma,
                             stretching_type=stretching, order=6)
    rng = np.random.default_rng(7)
    zq = rng.uniform(0.0, Lz, 4000)
    zq = np.concatenate([zq, z_f.numpy(), z_c.numpy()[1:-1]])
    kf = adv._face_coord(torch.from_numpy(zq)).numpy()
    d = {
        "gamma": gamma, "nz": nz, "Lz": Lz, "stretching": stretching,
        "z_f": z_f.numpy(), "z_c": z_c.numpy(), "dz_f": dz_f.numpy(), "dz_c": dz_c.numpy(),
        "tanh_g": adv._tanh_g, "z_lo": adv.z_lo, "z_hi": adv.z_hi,
        "zq": zq, "kf": kf,
    }
    for comp in "uvw":
        for o in (2, 4, 6):
            d[f"dinv_{comp}{o}"] = semilag._z_denominator_table(adv.spec[comp].znodes, o).numpy()
    np.savez_compressed(os.path.join(HERE, f"grid_{name}.npz"), **d)


def gen_advect(ref, name, order, stretching="symmetric", clamp=False, seed=1, traj_order=2):
    torch, semilag, solver, utils, projection, operators = ref
    nx, ny, nz = 12, 10, 16
    Lx, Ly, Lz, gamma = 4 * math.pi, 2 * math.pi, 2.0, 1.5
    dx, dy = Lx / nx, Ly / ny
    z_f, z_c, _, _ = utils.generate_grid(gamma, nz, Lz, stretching_type=stretching)
    zf, zc = z_f.numpy(), z_c.numpy()
    rng = np.random.default_rng(seed)
    u, v, w = make_fields(rng, nx, ny, nz, dx, dy, Lx, Ly, Lz, zf, zc)
    tu, tv, tw = (torch.from_numpy(a.copy()) for a in (u, v, w))
    solver.apply_bc_all(tu, tv, tw, "dirichlet")
    u, v, w = tu.numpy().copy(), tv.numpy().copy(), tw.numpy().copy()
    env = lambda X, Y, Z: 0.0 * X  # noqa: E731
    base = {"u": lambda X, Y, Z: 1.0 + 0.4 * np.sin(math.pi * Z / Lz) + env(X, Y, Z)}
    wamp = 2.5 if clamp else 0.6
    um, vm, wm = make_fields(rng, nx, ny, nz, dx, dy, Lx, Ly, Lz, zf, zc, base=base,
                             amp=0.6, wamp=wamp)
    tum, tvm, twm = (torch.from_numpy(a.copy()) for a in (um, vm, wm))
    solver.apply_bc_all(tum, tvm, twm, "dirichlet")
    um, vm, wm = tum.numpy().copy(), tvm.numpy().copy(), twm.numpy().copy()
    dt = 1.6 if not clamp else 0.9

    adv = semilag.SLAdvector(nx, ny, nz, dx, dy, Lx, Ly, Lz, z_f, z_c, gamma,
                             stretching_type=stretching, order=order, traj_order=traj_order)
    ts = [torch.from_numpy(a.copy()) for a in (u, v, w, um, vm, wm)]
    ustar, vstar, wstar = adv.advect(*ts, torch.tensor(dt))
    d = {
        "nx": nx, "ny": ny, "nz": nz, "Lx": Lx, "Ly": Ly, "Lz": Lz, "gamma": gamma,
        "stretching": stretching, "order": order, "dt": dt, "traj_order": traj_order,
        "z_f": zf, "z_c": zc,
        "u": u, "v": v, "w": w, "um": um, "vm": vm, "wm": wm,
        "ustar": ustar.numpy().copy(), "vstar": vstar.numpy().copy(),
        "wstar": wstar.numpy().copy(), "n_clamped": int(adv.n_clamped_last.item()),
    }
    # feet and stencil base indices per component (mids already in mbuf)
    for comp in "uvw":
        spec = adv.spec[comp]
        xd, yd, zd = adv.departure_coords(comp, torch.tensor(dt))
        iw = adv.compute_departure(comp, torch.tensor(dt))
        P = order
        d[f"xd_{comp}"] = xd.reshape(-1).numpy().copy()
        d[f"yd_{comp}"] = yd.reshape(-1).numpy().copy()
        d[f"zd_{comp}"] = zd.reshape(-1).numpy().copy()
        # base (offset P/2-1 -> index 0 of the stencil) x/y node indices
        d[f"ix_{comp}"] = (iw.ix_lin[:, 0] // (spec.NY * spec.NZ)).to(torch.int32).numpy()
        d[f"iy_{comp}"] = (iw.iy_lin[:, 0] // spec.NZ).to(torch.int32).numpy()
        d[f"k0_{comp}"] = iw.kz0.to(torch.int32).numpy()
        del P
    np.savez_compressed(os.path.join(HERE, f"advect_{name}.npz"), **d)


def gen_solve(ref):
    torch, semilag, solver, utils, projection, operators = ref
    nx, ny, nz = 12, 10, 16
    Lx, Ly, Lz, gamma = 4 * math.pi, 2 * math.pi, 2.0, 1.5
    dx, dy = Lx / nx, Ly / ny
    rng = np.random.default_rng(11)
    d = {"nx": nx, "ny": ny, "nz": nz, "Lx": Lx, "Ly": Ly, "Lz": Lz, "gamma": gamma}
    for stretching, top in (("symmetric", "dirichlet"), ("bottom", "neumann")):
        tag = "sym" if stretching == "symmetric" else "bot"
        z_f, z_c, dz_f, dz_c = utils.generate_grid(gamma, nz, Lz, stretching_type=stretching)
        fft = projection.initialize_fft_solver(nx, ny, nz, dx, dy, dz_c, dz_f,
                                               top_wall_bc_type=top)
        div = rng.normal(size=(nx, ny, nz))
        div -= div.mean()
        p = projection.solve_poisson_fft(torch.from_numpy(div.copy()), fft)
        d[f"{tag}_z_f"] = z_f.numpy()
        d[f"{tag}_z_c"] = z_c.numpy()
        d[f"{tag}_div"] = div
        d[f"{tag}_p"] = p.numpy().copy()
        u = rng.normal(size=(nx + 1, ny + 2, nz + 2))
        v = rng.normal(size=(nx + 2, ny + 1, nz + 2))
        w = rng.normal(size=(nx + 2, ny + 2, nz + 1))
        d[f"{tag}_u"], d[f"{tag}_v"], d[f"{tag}_w"] = u, v, w
        nu = 1.0 / 50.0
        for theta in (1.0, 0.5):
            dt = 0.3
            su = operators.solve_implicit_diffusion_u(
                torch.from_numpy(u.copy()), dt, nx, ny, nz, dz_c, dz_f, nu,
                theta=theta, top_wall_bc_type=top)
            sv = operators.solve_implicit_diffusion_v(
                torch.from_numpy(v.copy()), dt, nx, ny, nz, dz_c, dz_f, nu,
                theta=theta, top_wall_bc_type=top)
            sw = operators.solve_implicit_diffusion_w(
                torch.from_numpy(w.copy()), dt, nx, ny, nz, dz_c, dz_f, nu, theta=theta)
            th = "t1" if theta == 1.0 else "t05"
            d[f"{tag}_{th}_u"] = su.numpy().copy()
            d[f"{tag}_{th}_v"] = sv.numpy().copy()
            d[f"{tag}_{th}_w"] = sw.numpy().copy()
        d[f"{tag}_nu"] = nu
        d[f"{tag}_dt"] = 0.3
        d[f"{tag}_top"] = top
    np.savez_compressed(os.path.join(HERE, "solve_small.npz"), **d)


def gen_step(ref, name, order, stretching="symmetric", top="dirichlet", n_steps=4, seed=3,
             sl_extra=None):
    torch, semilag, solver_mod, utils, projection, operators = ref
    nx, ny, nz = 12, 10, 16
    Lx, Ly, Lz, gamma = 4 * math.pi, 2 * math.pi, 2.0, 1.5
    Re, dt = 2800.0, 0.8
    import yaml

    with tempfile.TemporaryDirectory() as tmp:
        cfg = {
            "grid": {"nx": nx, "ny": ny, "nz": nz},
            "domain": {"Lx": Lx, "Ly": Ly, "Lz": Lz, "stretching_type": stretching},
            "flow": {"Re": Re, "Re_tau": 180.0, "U_bulk": 1.0, "gamma": gamma},
            "boundary_conditions": {"top_wall": {"type": top}},
            "initialization": {"type": "parabolic", "perturbation_intensity": 0.0},
            "solver": {"type": "fft"},
            "time": {"dt": dt, "n_steps": 10, "t_max": 1e9, "CFL_target": 3.0,
                     "dt_update_interval": 0, "dt_max": 10.0, "dt_min": 1e-6},
            "compute": {"device": "cpu"},
            "output": {"results_folder": os.path.join(tmp, "results"), "n_out": 10**8,
                       "n_save": 10**8},
            "advection": {"scheme": "sl"},
            "sl": {"interp_order": order, "n_traj_iters": 2, "interp_dtype": "fp64",
                   **(sl_extra or {})},
            "statistics": {"n_stats": 0},
        }
        path = os.path.join(tmp, "c.yaml")
        with open(path, "w") as f:
            yaml.safe_dump(cfg, f)
        s = solver_mod.SLChannelFlow(config_file=path)

        zf, zc = s.z_f.numpy(), s.z_c.numpy()
        dx, dy = s.dx, s.dy
        rng = np.random.default_rng(seed)
        base = {"u": lambda X, Y, Z: 1.5 * (1.0 - (Z - 1.0) ** 2)}
        u, v, w = make_fields(rng, nx, ny, nz, dx, dy, Lx, Ly, Lz, zf, zc, base=base,
                              amp=0.25, wamp=0.15)
        tu, tv, tw = (torch.from_numpy(a.copy()) for a in (u, v, w))
        solver_mod.apply_bc_all(tu, tv, tw, top)
        # same projection as the reference's initial-field projection (dt=1)
        div = utils.compute_divergence(tu, tv, tw, nx, ny, nz, dx, dy, s.dz_f)
        p = projection.solve_poisson_fft(div, s.fft_data).clone()
        tu, tv, tw = projection.project_velocity(tu, tv, tw, p, nx, ny, nz, dx, dy,
                                                 s.dz_c, s.dz_f, 1.0)
        solver_mod.apply_bc_all(tu, tv, tw, top)
        s.u, s.v, s.w, s.p = tu, tv, tw, p
        s.forcing = 0.0
        d = {
            "nx": nx, "ny": ny, "nz": nz, "Lx": Lx, "Ly": Ly, "Lz": Lz, "gamma": gamma,
            "Re": Re, "dt": dt, "order": order, "stretching": stretching, "top": top,
            "bdf2_pressure": (sl_extra or {}).get("bdf2_pressure", "noninc"),
            "bdf2_xy_rhs": (sl_extra or {}).get("bdf2_xy_rhs", "extrap"),
            "traj_order": (sl_extra or {}).get("traj_interp_order", 2),
            "z_f": zf, "z_c": zc,
            "u0": tu.numpy().copy(), "v0": tv.numpy().copy(), "w0": tw.numpy().copy(),
            "p0": p.numpy().copy(),
        }
        ub, fo, ncl = [], [], []
        for n in range(n_steps):
            u_b, forcing = s.step_sl_bdf2(dt)
            ub.append(float(u_b))
            fo.append(float(forcing))
            ncl.append(int(s.sl.n_clamped_last.item()))
            if n in (0, n_steps - 1):
                d[f"u_{n + 1}"] = s.u.numpy().copy()
                d[f"v_{n + 1}"] = s.v.numpy().copy()
                d[f"w_{n + 1}"] = s.w.numpy().copy()
                d[f"p_{n + 1}"] = s.p.numpy().copy()
        d["u_bulk"] = np.array(ub)
        d["forcing"] = np.array(fo)
        d["n_clamped"] = np.array(ncl)
        d["n_steps"] = n_steps
    np.savez_compressed(os.path.join(HERE, f"step_{name}.npz"), **d)


def _stepper_config(tmp, nx, ny, nz, Lx, Ly, Lz, gamma, Re, dt, stretching, top, advection,
                    sl=None):
    import yaml

    cfg = {
        "grid": {"nx": nx, "ny": ny, "nz": nz},
        "domain": {"Lx": Lx, "Ly": Ly, "Lz": Lz, "stretching_type": stretching},
        "flow": {"Re": Re, "Re_tau": 180.0, "U_bulk": 1.0, "gamma": gamma},
        "boundary_conditions": {"top_wall": {"type": top}},
        "initialization": {"type": "parabolic", "perturbation_intensity": 0.0},
        "solver": {"type": "fft"},
        "time": {"dt": dt, "n_steps": 10, "t_max": 1e9, "CFL_target": 3.0,
                 "dt_update_interval": 0, "dt_max": 10.0, "dt_min": 1e-6},
        "compute": {"device": "cpu"},
        "output": {"results_folder": os.path.join(tmp, "results"), "n_out": 10**8,
                   "n_save": 10**8},
        "advection": {"scheme": advection},
        "statistics": {"n_stats": 0},
    }
    if sl is not None:
        cfg["sl"] = sl
    path = os.path.join(tmp, "c.yaml")
    with open(path, "w") as f:
        yaml.safe_dump(cfg, f)
    return path


def gen_imex(ref):
    torch, semilag, solver_mod, utils, projection, operators = ref
    nx, ny, nz = 12, 10, 16
    Lx, Ly, Lz, gamma = 4 * math.pi, 2 * math.pi, 2.0, 1.5
    dx, dy = Lx / nx, Ly / ny
    d = {"nx": nx, "ny": ny, "nz": nz, "Lx": Lx, "Ly": Ly, "Lz": Lz, "gamma": gamma}
    rng = np.random.default_rng(21)
    nu = 1.0 / 400.0
    for stretching, top in (("symmetric", "dirichlet"), ("bottom", "neumann")):
        tag = "sym" if stretching == "symmetric" else "bot"
        z_f, z_c, dz_f, dz_c = utils.generate_grid(gamma, nz, Lz, stretching_type=stretching)
        zf, zc = z_f.numpy(), z_c.numpy()
        base = {"u": lambda X, Y, Z: 1.5 * (1.0 - (Z - 1.0) ** 2)}
        u, v, w = make_fields(rng, nx, ny, nz, dx, dy, Lx, Ly, Lz, zf, zc, base=base,
                              amp=0.3, wamp=0.2)
        tu, tv, tw = (torch.from_numpy(a.copy()) for a in (u, v, w))
        solver_mod.apply_bc_all(tu, tv, tw, top)
        ru, rv, rw = operators.compute_momentum_rhs_fused_imex(
            tu, tv, tw, nx, ny, nz, dx, dy, dz_c, dz_f, nu)
        d[f"{tag}_top"] = top
        d[f"{tag}_nu"] = nu
        d[f"{tag}_z_f"], d[f"{tag}_z_c"] = zf, zc
        d[f"{tag}_u"], d[f"{tag}_v"], d[f"{tag}_w"] = (t.numpy().copy() for t in (tu, tv, tw))
        d[f"{tag}_ru"], d[f"{tag}_rv"], d[f"{tag}_rw"] = (t.numpy().copy() for t in (ru, rv, rw))

    # three IMEX steps from a seeded projected state
    Re, dt, n_steps = 400.0, 0.05, 3
    with tempfile.TemporaryDirectory() as tmp:
        path = _stepper_config(tmp, nx, ny, nz, Lx, Ly, Lz, gamma, Re, dt, "symmetric",
                               "dirichlet", "eulerian")
        s = solver_mod.SLChannelFlow(config_file=path)
        # the right-hand side the reference computes on a GPU (solver.py:1027-1039)
        s.compute_momentum_rhs_explicit_imex = lambda: operators.compute_momentum_rhs_fused_imex(
            s.u, s.v, s.w, s.nx, s.ny, s.nz, s.dx, s.dy, s.dz_c, s.dz_f, s.nu)
        zf, zc = s.z_f.numpy(), s.z_c.numpy()
        base = {"u": lambda X, Y, Z: 1.5 * (1.0 - (Z - 1.0) ** 2)}
        u, v, w = make_fields(rng, nx, ny, nz, s.dx, s.dy, Lx, Ly, Lz, zf, zc, base=base,
                              amp=0.25, wamp=0.15)
        tu, tv, tw = (torch.from_numpy(a.copy()) for a in (u, v, w))
        solver_mod.apply_bc_all(tu, tv, tw, "dirichlet")
        div = utils.compute_divergence(tu, tv, tw, nx, ny, nz, s.dx, s.dy, s.dz_f)
        p = projection.solve_poisson_fft(div, s.fft_data).clone()
        tu, tv, tw = projection.project_velocity(tu, tv, tw, p, nx, ny, nz, s.dx, s.dy,
                                                 s.dz_c, s.dz_f, 1.0)
        solver_mod.apply_bc_all(tu, tv, tw, "dirichlet")
        s.u, s.v, s.w, s.p = tu, tv, tw, p
        s.forcing = 0.0
        d.update({"step_Re": Re, "step_dt": dt, "step_n": n_steps, "step_z_f": zf,
                  "step_z_c": zc, "step_u0": tu.numpy().copy(), "step_v0": tv.numpy().copy(),
                  "step_w0": tw.numpy().copy(), "step_p0": p.numpy().copy()})
        ub, fo = [], []
        for n in range(n_steps):
            u_b, forcing = s.step_imex(dt)
            ub.append(float(u_b))
            fo.append(float(forcing))
            for k in "uvwp":
                d[f"step_{k}_{n + 1}"] = getattr(s, k).numpy().copy()
        d["step_u_bulk"] = np.array(ub)
        d["step_forcing"] = np.array(fo)
    np.savez_compressed(os.path.join(HERE, "imex_small.npz"), **d)


def gen_stats(ref):
    torch, semilag, solver_mod, utils, projection, operators = ref
    import contextlib
    import io

    from slchannel import turbstats

    nx, ny, nz = 16, 12, 24
    Lx, Ly, Lz, gamma = 4 * math.pi, 2 * math.pi, 2.0, 1.5
    dx, dy = Lx / nx, Ly / ny
    nu = 1.0 / 3000.0
    z_f, z_c, dz_f, dz_c = utils.generate_grid(gamma, nz, Lz, stretching_type="symmetric")
    zf, zc = z_f.numpy(), z_c.numpy()
    with contextlib.redirect_stdout(io.StringIO()):
        ts = turbstats.TurbulenceStats(nx, ny, nz, Lx, Ly, Lz, z_c, z_f, dz_c, dz_f, dx, dy,
                                       nu, 180.0, z_plus_target=15.0, device="cpu",
                                       top_wall_bc_type="dirichlet",
                                       spectra_z_planes=[5.0, 40.0])
    rng = np.random.default_rng(31)
    d = {"nx": nx, "ny": ny, "nz": nz, "Lx": Lx, "Ly": Ly, "Lz": Lz, "gamma": gamma, "nu": nu,
         "z_f": zf, "z_c": zc, "k_bot": ts.k_bot, "k_top": ts.k_top,
         "extra_k": np.array([[pl["k_bot"], pl["k_top"]] for pl in ts.extra_planes])}
    base = {"u": lambda X, Y, Z: 1.5 * (1.0 - (Z - 1.0) ** 2),
            "v": lambda X, Y, Z: 0.05 + 0.0 * Z}
    for n in range(3):
        u, v, w = make_fields(rng, nx, ny, nz, dx, dy, Lx, Ly, Lz, zf, zc, base=base,
                              amp=0.3, wamp=0.2)
        tu, tv, tw = (torch.from_numpy(a.copy()) for a in (u, v, w))
        solver_mod.apply_bc_all(tu, tv, tw, "dirichlet")
        d[f"u{n}"], d[f"v{n}"], d[f"w{n}"] = (t.numpy().copy() for t in (tu, tv, tw))
        ts.accumulate_statistics(tu, tv, tw, 0.05)
    for name in ("U_sum", "uu_sum", "vv_sum", "ww_sum", "uw_sum", "Vm_sum", "Wm_sum",
                 "Um2_sum", "Vm2_sum", "Wm2_sum", "UmWm_sum", "E_uu_2d_sum", "E_vv_2d_sum",
                 "E_ww_2d_sum", "E_uw_2d_sum"):
        d[name] = getattr(ts, name).numpy().copy()
    for p, pl in enumerate(ts.extra_planes):
        for c in ("uu", "vv", "ww", "uw"):
            d[f"extra{p}_{c}"] = pl["sums"][c].numpy().copy()
    with contextlib.redirect_stdout(io.StringIO()):
        fin = ts.finalize_statistics()
    for k in ("U_mean", "uu_mean", "vv_mean", "ww_mean", "uw_mean", "E_uu_2d", "E_vv_2d",
              "E_ww_2d", "E_uw_2d", "u_tau", "n_samples"):
        d[f"fin_{k}"] = np.asarray(fin[k])
    for pl in ts.extra_planes:
        for c in ("uu", "vv", "ww", "uw"):
            key = f"E_{c}_2d_z{pl['zp']:.0f}"
            d[f"fin_{key}"] = np.asarray(fin[key])
    np.savez_compressed(os.path.join(HERE, "stats_small.npz"), **d)


def gen_diag(ref):
    torch, semilag, solver_mod, utils, projection, operators = ref
    nx, ny, nz = 12, 10, 16
    Lx, Ly, Lz, gamma = 4 * math.pi, 2 * math.pi, 2.0, 1.5
    dx, dy = Lx / nx, Ly / ny
    z_f, z_c, dz_f, dz_c = utils.generate_grid(gamma, nz, Lz, stretching_type="symmetric")
    zf, zc = z_f.numpy(), z_c.numpy()
    rng = np.random.default_rng(41)
    base = {"u": lambda X, Y, Z: 1.5 * (1.0 - (Z - 1.0) ** 2)}
    u, v, w = make_fields(rng, nx, ny, nz, dx, dy, Lx, Ly, Lz, zf, zc, base=base, amp=0.4,
                          wamp=0.3)
    tu, tv, tw = (torch.from_numpy(a.copy()) for a in (u, v, w))
    solver_mod.apply_bc_all(tu, tv, tw, "dirichlet")
    nu = 1.0 / 2800.0
    d = {"nx": nx, "ny": ny, "nz": nz, "Lx": Lx, "Ly": Ly, "Lz": Lz, "gamma": gamma, "nu": nu,
         "z_f": zf, "z_c": zc, "u": tu.numpy().copy(), "v": tv.numpy().copy(),
         "w": tw.numpy().copy()}
    d["u_tau_dirichlet"] = float(utils.compute_u_tau(tu, z_c, nu, "dirichlet"))
    d["u_tau_neumann"] = float(utils.compute_u_tau(tu, z_c, nu, "neumann"))
    d["dti"] = float(operators.compute_cfl_fused(tu, tv, tw, nx, ny, nz, dx, dy, dz_f, dz_c))
    np.savez_compressed(os.path.join(HERE, "diag_small.npz"), **d)


def main():
    ref_path = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    only = set(sys.argv[2:])
    ref = _import_reference(ref_path)
    cases = [
        ("grid_sym_g15_n16", lambda: gen_grid(ref, "sym_g15_n16", 1.5, 16, 2.0, "symmetric")),
        ("grid_bot_g18_n24", lambda: gen_grid(ref, "bot_g18_n24", 1.8, 24, 2.0, "bottom")),
        ("grid_sym_g15_n64", lambda: gen_grid(ref, "sym_g15_n64", 1.5, 64, 2.0, "symmetric")),
        ("advect_o4", lambda: gen_advect(ref, "o4", 4)),
        ("advect_o6", lambda: gen_advect(ref, "o6", 6)),
        ("advect_o6_bottom", lambda: gen_advect(ref, "o6_bottom", 6, stretching="bottom", seed=2)),
        ("advect_o4_clamp", lambda: gen_advect(ref, "o4_clamp", 4, clamp=True, seed=5)),
        ("solve_small", lambda: gen_solve(ref)),
        ("step_o6", lambda: gen_step(ref, "o6", 6)),
        ("step_o4_bottom_neumann", lambda: gen_step(ref, "o4_bottom_neumann", 4,
                                                    stretching="bottom", top="neumann", seed=4)),
        ("step_o4_inc", lambda: gen_step(ref, "o4_inc", 4, seed=6,
                                         sl_extra={"bdf2_pressure": "inc"})),
        ("step_o6_lagged", lambda: gen_step(ref, "o6_lagged", 6, seed=7,
                                            sl_extra={"bdf2_xy_rhs": "lagged"})),
        ("advect_o6_traj4", lambda: gen_advect(ref, "o6_traj4", 6, seed=8, traj_order=4)),
        ("step_o4_traj4", lambda: gen_step(ref, "o4_traj4", 4, seed=9,
                                           sl_extra={"traj_interp_order": 4})),
        ("imex_small", lambda: gen_imex(ref)),
        ("stats_small", lambda: gen_stats(ref)),
        ("diag_small", lambda: gen_diag(ref)),
    ]
    for name, make in cases:
        if not only or name in only:
            make()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
