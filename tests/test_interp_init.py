"""Restart onto another grid: initialization.type 'interpolate'
(initflow.py:47-120, solver.py:351-401 and 599-618).

The reference's tests hold no vectors for this path, so the oracle
(oracle.interpolate_fields and its two helpers) is pinned on the CPU by what
the mathematics fixes:
* the 1-D weights reproduce numpy's piecewise-linear interpolation, with
  constant extrapolation outside the nodes;
* the tensor-product sum reproduces any trilinear polynomial exactly inside
  the source box;
* a field interpolated onto its own grid comes back unchanged, ghosts
  included (a half-cell offset or a swapped face/centre table fails this);
* from a coarser source of twice the box, a field linear in x, y and z is
  reproduced at every target node whose bracketing source nodes are
  interior (checks the proportional x/y mapping and the z lookup);
* from a source twice as tall, the target reads the lower part (a field
  linear in z, source_half 'lower').

GPU: the device interpolation (slc_interp_trilinear through
checkpoint.interpolated_fields) equals the oracle bit for bit (same
operations in the same order, no contraction), and the solver's interpolate
start (boundary fill, u rescaled to U_bulk, projection) equals the oracle's
sequence to 1e-12 with u_bulk = U_bulk and max|div| < 1e-11.
"""

import math

import numpy as np
import pytest
import torch

import oracle.oracle as orc
from slchannel_amd.checkpoint import save_flow_fields

LX, LY = 2 * math.pi, math.pi


def _zgrid(nz, Lz, gamma=1.5):
    z_f, z_c, _, _ = orc.generate_grid(gamma, nz, Lz, "symmetric")
    return np.asarray(z_f, dtype=np.float64), np.asarray(z_c, dtype=np.float64)


def _coords(n, L):
    d = L / n
    return np.arange(n + 1) * d, (np.arange(n + 2) - 0.5) * d


def _source(nx, ny, nz, Lz, seed):
    rng = np.random.default_rng(seed)
    z_f, z_c = _zgrid(nz, Lz)
    return (rng.standard_normal((nx + 1, ny + 2, nz + 2)),
            rng.standard_normal((nx + 2, ny + 1, nz + 2)),
            rng.standard_normal((nx + 2, ny + 2, nz + 1)), z_c, z_f)


def test_interp1d_weights_is_piecewise_linear(check):
    rng = np.random.default_rng(3)
    nodes = np.sort(rng.uniform(-1.0, 3.0, 40))
    f = rng.standard_normal(40)
    q = np.concatenate([rng.uniform(-2.0, 4.0, 500), nodes, [nodes[0] - 1.0, nodes[-1] + 1.0]])
    k0, t = orc.interp1d_weights(q, nodes)
    got = (1.0 - t) * f[k0] + t * f[k0 + 1]
    err = np.abs(got - np.interp(q, nodes, f)).max()
    check("(1-t) f[k0] + t f[k0+1] == np.interp (constant outside)", err < 1e-13, f"{err:.1e}")
    check("indices within [0, len-2], weights in [0, 1]",
          k0.min() >= 0 and k0.max() <= len(nodes) - 2 and t.min() >= 0 and t.max() <= 1, "")


def test_trilinear_tensor_reproduces_trilinear_polynomials(check):
    rng = np.random.default_rng(4)
    xn, yn = np.sort(rng.uniform(0, 2, 9)), np.sort(rng.uniform(-1, 1, 7))
    zn = np.sort(rng.uniform(0, 1, 11))
    c = rng.standard_normal(8)

    def f(x, y, z):
        return (c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * y + c[5] * y * z
                + c[6] * x * z + c[7] * x * y * z)

    field = f(xn[:, None, None], yn[None, :, None], zn[None, None, :])
    xq = rng.uniform(xn[0], xn[-1], 13)
    yq = rng.uniform(yn[0], yn[-1], 6)
    zq = rng.uniform(zn[0], zn[-1], 17)
    got = orc.trilinear_tensor(field, xn, yn, zn, xq, yq, zq)
    want = f(xq[:, None, None], yq[None, :, None], zq[None, None, :])
    err = np.abs(got - want).max()
    check("trilinear polynomial reproduced inside the box", err < 1e-12, f"{err:.1e}")
    # outside: the nearest face's value (constant extrapolation in each axis)
    out = orc.trilinear_tensor(field, xn, yn, zn, np.array([xn[0] - 5.0]), yq, zq)
    edge = f(xn[0], yq[:, None], zq[None, :])[None]
    err = np.abs(out - edge).max()
    check("below the first x node: the x = x_0 face", err < 1e-12, f"{err:.1e}")


def _refreshed(u, v, w):
    """The source's periodic ghost refresh of initflow.py:83-89."""
    u, v, w = u.copy(), v.copy(), w.copy()
    u[0] = u[-1]
    u[:, 0] = u[:, -2]
    u[:, -1] = u[:, 1]
    v[0] = v[-2]
    v[-1] = v[1]
    v[:, 0] = v[:, -1]
    w[0] = w[-2]
    w[-1] = w[1]
    w[:, 0] = w[:, -2]
    w[:, -1] = w[:, 1]
    return u, v, w


def test_same_grid_is_identity(check):
    nx, ny, nz, Lz = 8, 6, 10, 2.0
    u, v, w, z_c, z_f = _source(nx, ny, nz, Lz, 5)
    got = orc.interpolate_fields(u, v, w, z_c, z_f, LX, LY, nx, ny, nz, LX, LY, z_c, z_f)
    want = _refreshed(u, v, w)
    for name, a, b in zip("uvw", got, want):
        check(f"{name}: same grid gives the refreshed source exactly", np.array_equal(a, b),
              f"max diff {np.abs(a - b).max():.1e}")


def _linear(x, y, z):
    return 0.25 + 0.5 * x - 0.75 * y + 1.25 * z


def test_coarse_wide_source_reproduces_linear_field(check):
    """Source 6 x 5 x 8 over (2 Lx, 2 Ly); target 16 x 12 x 20 over (Lx, Ly):
    target node x maps to x * 2 in the source, so the expected value is the
    linear field at the mapped coordinates."""
    nxs, nys, nzs, Lz = 6, 5, 8, 2.0
    Lxs, Lys = 2 * LX, 2 * LY
    z_f_s, z_c_s = _zgrid(nzs, Lz)
    xfs, xcs = _coords(nxs, Lxs)
    yfs, ycs = _coords(nys, Lys)
    u_s = _linear(xfs[:, None, None], ycs[None, :, None], z_c_s[None, None, :])
    v_s = _linear(xcs[:, None, None], yfs[None, :, None], z_c_s[None, None, :])
    w_s = _linear(xcs[:, None, None], ycs[None, :, None], z_f_s[None, None, :])
    nx, ny, nz = 16, 12, 20
    z_f, z_c = _zgrid(nz, Lz, gamma=1.2)
    u, v, w = orc.interpolate_fields(u_s, v_s, w_s, z_c_s, z_f_s, Lxs, Lys, nx, ny, nz, LX, LY,
                                     z_c, z_f)
    xf, xc = _coords(nx, LX)
    yf, yc = _coords(ny, LY)
    xf, xc, yf, yc = 2 * xf, 2 * xc, 2 * yf, 2 * yc  # mapped onto the source box
    # interior brackets only: the refresh rewrites the source's periodic ghosts
    zi = (z_c > z_c_s[0]) & (z_c < z_c_s[-1])
    zfi = (z_f >= z_f_s[0]) & (z_f <= z_f_s[-1])
    for name, got, xq, yq, zq, zm, xs, ys in (
            ("u", u, xf, yc, z_c, zi, xfs, ycs), ("v", v, xc, yf, z_c, zi, xcs, yfs),
            ("w", w, xc, yc, z_f, zfi, xcs, ycs)):
        xm = (xq > xs[1]) & (xq < xs[-2])
        ym = (yq > ys[1]) & (yq < ys[-2])
        want = _linear(xq[:, None, None], yq[None, :, None], zq[None, None, :])
        sel = np.ix_(xm, ym, zm)
        err = np.abs(got[sel] - want[sel]).max()
        check(f"{name}: linear field at mapped nodes ({xm.sum()}x{ym.sum()}x{zm.sum()})",
              err < 1e-12 and xm.sum() > 4 and ym.sum() > 4, f"{err:.1e}")


def test_taller_source_gives_lower_part(check):
    nxs, nys, nzs = 8, 6, 24
    z_f_s, z_c_s = _zgrid(nzs, 2.0)
    xfs, xcs = _coords(nxs, LX)
    yfs, ycs = _coords(nys, LY)
    zfun = lambda z: 1.0 + 3.0 * z  # noqa: E731
    u_s = np.broadcast_to(zfun(z_c_s), (nxs + 1, nys + 2, nzs + 2)).copy()
    v_s = np.broadcast_to(zfun(z_c_s), (nxs + 2, nys + 1, nzs + 2)).copy()
    w_s = np.broadcast_to(zfun(z_f_s), (nxs + 2, nys + 2, nzs + 1)).copy()
    nz = 12
    z_f, z_c = np.linspace(0.0, 1.0, nz + 1), None
    z_c = np.concatenate([[-0.5 / nz], 0.5 * (z_f[1:] + z_f[:-1]), [1.0 + 0.5 / nz]])
    u, v, w = orc.interpolate_fields(u_s, v_s, w_s, z_c_s, z_f_s, LX, LY, nxs, nys, nz, LX, LY,
                                     z_c, z_f)
    err = max(np.abs(u[:, :, 1:-1] - zfun(z_c[1:-1])).max(),
              np.abs(v[:, :, 1:-1] - zfun(z_c[1:-1])).max(), np.abs(w - zfun(z_f)).max())
    check("open-channel target reads the source's lower half", err < 1e-12, f"{err:.1e}")


def _write_source(tmp_path, nx, ny, nz, Lx, Ly, Lz, seed):
    u, v, w, z_c, z_f = _source(nx, ny, nz, Lz, seed)
    u[:, :, 1:-1] += 1.0  # a mean flow, so the rescale is exercised
    p = np.zeros((nx + 2, ny + 2, nz + 2))
    save_flow_fields(*(torch.tensor(a) for a in (u, v, w, p, z_c, z_f)), Lx, Ly, 500, 12.5,
                     0.05, 0.002, str(tmp_path), "src.npz")
    return str(tmp_path / "src.npz"), (u, v, w, z_c, z_f)


@pytest.mark.gpu
def test_device_interpolation_equals_oracle(check, tmp_path):
    from slchannel_amd.checkpoint import interpolated_fields, load_flow_fields

    path, _ = _write_source(tmp_path, 12, 10, 16, 2 * LX, 1.5 * LY, 2.0, 7)
    src = load_flow_fields(path)
    for nx, ny, nz, Lz in ((20, 16, 24, 2.0), (7, 5, 9, 1.0)):
        z_f, z_c = _zgrid(nz, Lz, gamma=1.3)
        print(f"interpolated_fields onto {nx}x{ny}x{nz} Lz={Lz}", flush=True)
        got = interpolated_fields(path, nx, ny, nz, LX, LY, Lz, torch.tensor(z_c),
                                  torch.tensor(z_f), "cuda:0")
        want = orc.interpolate_fields(src["u"].numpy(), src["v"].numpy(), src["w"].numpy(),
                                      src["z_c"].numpy(), src["z_f"].numpy(), src["Lx"],
                                      src["Ly"], nx, ny, nz, LX, LY, z_c, z_f)
        for name, a, b in zip("uvw", got[:3], want):
            a = a.cpu().numpy()
            check(f"{nx}x{ny}x{nz} {name}: device == oracle bit for bit",
                  a.shape == b.shape and np.array_equal(a, b),
                  f"max diff {np.abs(a - b).max() if a.shape == b.shape else a.shape}")
        check(f"{nx}x{ny}x{nz} p is zero", not bool(got[3].any()), "")


@pytest.mark.gpu
def test_solver_interpolate_start(check, tmp_path):
    from slchannel_amd import SLChannelFlow
    from slchannel_amd.checkpoint import load_flow_fields

    path, _ = _write_source(tmp_path, 12, 10, 16, 2 * LX, 1.5 * LY, 2.0, 8)
    nx, ny, nz = 24, 16, 20
    cfg = {"grid": {"nx": nx, "ny": ny, "nz": nz},
           "domain": {"Lx": LX, "Ly": LY, "Lz": 2.0},
           "flow": {"Re": 2000.0, "U_bulk": 1.0, "gamma": 1.5},
           "time": {"dt": 0.01}, "sl": {"interp_order": 6},
           "initialization": {"type": "interpolate", "field_file": path}}
    print("SLChannelFlow interpolate start 24x16x20", flush=True)
    flow = SLChannelFlow(config=cfg)
    check("step and time reset", flow.current_step == 0 and flow.time == 0.0
          and flow.forcing == 0.0, f"{flow.current_step} {flow.time} {flow.forcing}")
    ub = flow.bulk_velocity().item()
    check("u_bulk == U_bulk", abs(ub - 1.0) < 1e-12, f"{ub!r}")
    div = flow.divergence().abs().max().item()
    check("max|div| < 1e-11", div < 1e-11, f"{div:.1e}")
    src = load_flow_fields(path)
    z_f, z_c = flow.z_f.numpy(), flow.z_c.numpy()
    co = orc.ChannelOracle(nx, ny, nz, LX, LY, 2.0, 1.5, 2000.0, order=6, z_f=z_f, z_c=z_c,
                           tanh_g=flow.plan.tanh_g)
    co.set_fields(*orc.interpolate_fields(src["u"].numpy(), src["v"].numpy(), src["w"].numpy(),
                                          src["z_c"].numpy(), src["z_f"].numpy(), src["Lx"],
                                          src["Ly"], nx, ny, nz, LX, LY, z_c, z_f))
    co.apply_bc_uvw()
    ub0 = orc.compute_bulk_velocity(co.u, co.cell_vol_ratio, co.total_volume)
    co.u *= co.U_bulk / ub0
    co.project_initial()
    worst = max(np.abs(getattr(flow, k).cpu().numpy() - getattr(co, k)).max()
                / max(np.abs(getattr(co, k)).max(), 1e-300) for k in "uvwp")
    check("fields == oracle's interpolate, fill, rescale, project", worst < 1e-12,
          f"max rel={worst:.1e}")
    u0 = flow.u.clone()
    flow.step_sl_bdf2(0.01)
    check("first step runs (BDF1 bootstrap) and stays finite",
          bool(torch.isfinite(flow.u).all()) and not torch.equal(flow.u, u0), "")


def test_taller_source_needs_lower_half(check, tmp_path):
    """initflow.py:78-79: only the lower part of a taller source is supported;
    another source_half is refused before any device work."""
    from slchannel_amd.checkpoint import interpolated_fields

    path, _ = _write_source(tmp_path, 6, 4, 8, LX, LY, 2.0, 9)
    z_f, z_c = _zgrid(4, 1.0)
    try:
        interpolated_fields(path, 6, 4, 4, LX, LY, 1.0, z_c, z_f, "cpu", source_half="upper")
        msg = None
    except ValueError as e:
        msg = str(e)
    check("source_half 'upper' with a taller source raises ValueError",
          msg is not None and "source_half" in msg, f"{msg}")
