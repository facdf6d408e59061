"""CPU oracle for the slChannel SL-BDF2 step -- TEST INFRASTRUCTURE ONLY.

A restatement of the reference's float64 eager path (gmcavallazzi/SLchannel),
used exclusively as the checker of the HIP layer: by ``tests/``, by
``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of ``bench.py``.
The product (``slchannel_amd``) never imports it.

Split:
* numpy (this file): grid, boundary conditions, explicit xy-diffusion, PCR and
  the theta-method z-solves, divergence, FFT-Poisson (numpy.fft / pocketfft),
  projection, bulk forcing and the BDF1/BDF2 step orchestration -- written
  vectorised in the same operation order as the reference;
* C (``sl_oracle.c``, built by ``oracle/Makefile``): departure feet and the
  Lagrange gathers, point by point, no FMA contraction.

Every function cites the reference lines it restates.  Parity status of each
function (what pins it) is listed in the repository's design notes; in short:
golden vectors written from the reference pin the grid, the feet, the stencil
indices, advect, both solves and full steps; analytic tests pin the rest.
"""

import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libslc_oracle.so")
_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(
            f"oracle C library missing ({_LIB_PATH}); build it with `make -C oracle`"
        )
    lib = ctypes.CDLL(_LIB_PATH)
    P = ctypes.POINTER
    dp = P(ctypes.c_double)
    ip = P(ctypes.c_int32)
    gp = P(_OrcGrid)
    lib.orc_departure.argtypes = [gp, ctypes.c_int, dp, dp, dp, ctypes.c_double, dp, dp, dp]
    lib.orc_departure.restype = None
    lib.orc_gather.argtypes = [gp, ctypes.c_int, dp, dp, dp, dp, dp, ip, ip, ip]
    lib.orc_gather.restype = ctypes.c_long
    lib.orc_sample.argtypes = [gp, ctypes.c_int, dp, ctypes.c_long, dp, dp, dp, dp]
    lib.orc_sample.restype = None
    lib.orc_interp.argtypes = [gp, ctypes.c_int, dp, ctypes.c_long, dp, dp, dp, dp]
    lib.orc_interp.restype = None
    lib.orc_locate_z.argtypes = [gp, ctypes.c_int, ctypes.c_long, dp, ip]
    lib.orc_locate_z.restype = None
    lib.orc_face_coord.argtypes = [gp, ctypes.c_double]
    lib.orc_face_coord.restype = ctypes.c_double
    _lib = lib
    return lib


class _OrcGrid(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("order", ctypes.c_int), ("n_iters", ctypes.c_int), ("symmetric", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("Lz", ctypes.c_double),
        ("gamma", ctypes.c_double), ("tanh_g", ctypes.c_double),
        ("z_lo", ctypes.c_double), ("z_hi", ctypes.c_double),
        ("z_c", ctypes.POINTER(ctypes.c_double)), ("z_f", ctypes.POINTER(ctypes.c_double)),
        ("dinv_c2", ctypes.POINTER(ctypes.c_double)), ("dinv_f2", ctypes.POINTER(ctypes.c_double)),
        ("dinv_cP", ctypes.POINTER(ctypes.c_double)), ("dinv_fP", ctypes.POINTER(ctypes.c_double)),
        ("traj_order", ctypes.c_int),
        ("dinv_cT", ctypes.POINTER(ctypes.c_double)), ("dinv_fT", ctypes.POINTER(ctypes.c_double)),
    ]


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _iptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


# ---------------------------------------------------------------------------
# Grid and tables
# ---------------------------------------------------------------------------


def generate_grid(gamma, nz, Lz, stretching_type="symmetric"):
    """utils.py:21-61 (tanh-stretched faces, mirrored ghost centres)."""
    k = np.linspace(0, nz, nz + 1)
    if stretching_type == "bottom":
        xi = k / nz
        z_f = Lz * (1.0 - np.tanh(gamma * (1.0 - xi)) / np.tanh(np.float64(gamma)))
    else:
        xi = (2 * k / nz) - 1
        z_f = 0.5 * Lz * (1 + np.tanh(gamma * xi) / np.tanh(np.float64(gamma)))
    z_c_inn = 0.5 * (z_f[:-1] + z_f[1:])
    z_c = np.concatenate([[-z_c_inn[0]], z_c_inn, [2 * z_f[-1] - z_c_inn[-1]]])
    dz_f = z_f[1:] - z_f[:-1]
    dz_c = z_c[1:] - z_c[:-1]
    return z_f, z_c, dz_f, dz_c


def z_denominator_table(znodes, order):
    """semilag.py:102-114: 1/prod_{l!=m}(z[b+m]-z[b+l]) for every stencil b."""
    nb = len(znodes) - order + 1
    zn = np.stack([znodes[b : b + order] for b in range(nb)])  # (nb, order)
    diff = zn[:, :, None] - zn[:, None, :] + np.eye(order)[None]
    prod = diff[:, :, 0].copy()
    for l in range(1, order):
        prod = prod * diff[:, :, l]
    return 1.0 / prod


def uniform_inv_denominators(order):
    """semilag.py:117-127."""
    offs = [m - (order // 2 - 1) for m in range(order)]
    inv = []
    for m in range(order):
        d = 1.0
        for l in range(order):
            if l != m:
                d *= offs[m] - offs[l]
        inv.append(1.0 / d)
    return inv


# ---------------------------------------------------------------------------
# Semi-Lagrangian advector (tables + C kernels)
# ---------------------------------------------------------------------------


class SLOracle:
    """Restates SLAdvector (semilag.py:130-671) on the fp64 eager path."""

    def __init__(self, nx, ny, nz, dx, dy, Lx, Ly, Lz, z_f, z_c, gamma,
                 stretching_type="symmetric", order=4, n_traj_iters=2, traj_order=2):
        if order not in (4, 6):
            raise ValueError(f"order must be 4 or 6, got {order}")
        if traj_order not in (2, 4):  # semilag.py:206-209
            raise ValueError(
                f"sl.traj_interp_order must be 2 (trilinear) or 4 (tricubic), got {traj_order}")
        self.traj_order = traj_order
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dx, self.dy, self.Lx, self.Ly, self.Lz = dx, dy, Lx, Ly, Lz
        self.gamma = float(gamma)
        self.order = order
        self.stretching_type = stretching_type
        self.z_f = np.ascontiguousarray(z_f, dtype=np.float64)
        self.z_c = np.ascontiguousarray(z_c, dtype=np.float64)
        self.tanh_g = math.tanh(self.gamma)  # semilag.py:240-242 (torch.tanh of a scalar)
        eps_z = 1e-12 * Lz  # semilag.py:245-247
        self.z_lo = float(self.z_f[0] + eps_z)
        self.z_hi = float(self.z_f[-1] - eps_z)
        self.dinv = {
            ("c", 2): z_denominator_table(self.z_c, 2),
            ("f", 2): z_denominator_table(self.z_f, 2),
            ("c", order): z_denominator_table(self.z_c, order),
            ("f", order): z_denominator_table(self.z_f, order),
            ("c", traj_order): z_denominator_table(self.z_c, traj_order),
            ("f", traj_order): z_denominator_table(self.z_f, traj_order),
        }
        self._g = _OrcGrid(
            nx, ny, nz, order, n_traj_iters, 1 if stretching_type == "symmetric" else 0,
            dx, dy, Lz, self.gamma, self.tanh_g, self.z_lo, self.z_hi,
            _dptr(self.z_c), _dptr(self.z_f),
            _dptr(self.dinv[("c", 2)]), _dptr(self.dinv[("f", 2)]),
            _dptr(self.dinv[("c", order)]), _dptr(self.dinv[("f", order)]),
            traj_order, _dptr(self.dinv[("c", traj_order)]), _dptr(self.dinv[("f", traj_order)]),
        )
        self.n_clamped_last = 0

    def set_tanh_g(self, tanh_g):
        """Use a tanh(gamma) computed elsewhere (keeps tables bit-identical)."""
        self.tanh_g = float(tanh_g)
        self._g.tanh_g = self.tanh_g

    def n_arrival(self, comp):
        nzi = self.nz - 1 if comp == "w" else self.nz
        return self.nx * self.ny * nzi

    def arrival_shape(self, comp):
        return (self.nx, self.ny, self.nz - 1 if comp == "w" else self.nz)

    def face_coord(self, z):
        lib = _load()
        return np.array([lib.orc_face_coord(ctypes.byref(self._g), float(v)) for v in z])

    def locate_z(self, comp, z):
        """semilag.py:413-420 (z clamped into [z_lo, z_hi] first)."""
        lib = _load()
        z = np.ascontiguousarray(z, dtype=np.float64).ravel()
        out = np.empty(z.size, dtype=np.int32)
        lib.orc_locate_z(ctypes.byref(self._g), "uvw".index(comp), z.size, _dptr(z), _iptr(out))
        return out

    def departure(self, comp, um, vm, wm, dt):
        """semilag.py:534-552: feet (x, y unwrapped; z not yet clamped)."""
        lib = _load()
        n = self.n_arrival(comp)
        xd, yd, zd = (np.empty(n) for _ in range(3))
        um, vm, wm = (np.ascontiguousarray(a, dtype=np.float64) for a in (um, vm, wm))
        lib.orc_departure(ctypes.byref(self._g), "uvw".index(comp), _dptr(um), _dptr(vm),
                          _dptr(wm), float(dt), _dptr(xd), _dptr(yd), _dptr(zd))
        return xd, yd, zd

    def gather(self, comp, field, xd, yd, zd, want_index=False):
        """semilag.py:471-510: clamp+count, order-P stencil, gather."""
        lib = _load()
        n = self.n_arrival(comp)
        out = np.empty(n)
        field = np.ascontiguousarray(field, dtype=np.float64)
        ix = iy = k0 = None
        if want_index:
            ix, iy, k0 = (np.empty(n, dtype=np.int32) for _ in range(3))
        ncl = lib.orc_gather(ctypes.byref(self._g), "uvw".index(comp), _dptr(field),
                             _dptr(xd), _dptr(yd), _dptr(zd), _dptr(out),
                             _iptr(ix), _iptr(iy), _iptr(k0))
        if want_index:
            return out, int(ncl), (ix, iy, k0)
        return out, int(ncl)

    def sample(self, comp, field, x, y, z):
        """semilag.py:512-528 (trajectory sampling at traj_order: 2 trilinear,
        4 tricubic)."""
        lib = _load()
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (x, y, z))
        out = np.empty(x.size)
        field = np.ascontiguousarray(field, dtype=np.float64)
        lib.orc_sample(ctypes.byref(self._g), "uvw".index(comp), _dptr(field), x.size,
                       _dptr(x), _dptr(y), _dptr(z), _dptr(out))
        return out

    def interp(self, comp, field, x, y, z):
        """Order-P interpolation at points (semilag.py:460-510)."""
        lib = _load()
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (x, y, z))
        out = np.empty(x.size)
        field = np.ascontiguousarray(field, dtype=np.float64)
        lib.orc_interp(ctypes.byref(self._g), "uvw".index(comp), _dptr(field), x.size,
                       _dptr(x), _dptr(y), _dptr(z), _dptr(out))
        return out

    def advect(self, u, v, w, u_mid, v_mid, w_mid, dt, out=None):
        """semilag.py:584-671: returns ghost-shaped (ustar, vstar, wstar) with
        interiors set (w wall faces untouched); n_clamped in n_clamped_last."""
        nx, ny, nz = self.nx, self.ny, self.nz
        if out is None:
            out = (np.zeros((nx + 1, ny + 2, nz + 2)), np.zeros((nx + 2, ny + 1, nz + 2)),
                   np.zeros((nx + 2, ny + 2, nz + 1)))
        fields = {"u": u, "v": v, "w": w}
        ncl = 0
        for ic, comp in enumerate("uvw"):
            xd, yd, zd = self.departure(comp, u_mid, v_mid, w_mid, dt)
            vals, c = self.gather(comp, fields[comp], xd, yd, zd)
            ncl += c
            shape = self.arrival_shape(comp)
            if comp == "w":
                out[ic][1 : nx + 1, 1 : ny + 1, 1:nz] = vals.reshape(shape)
            else:
                out[ic][1 : nx + 1, 1 : ny + 1, 1 : nz + 1] = vals.reshape(shape)
        self.n_clamped_last = ncl
        return out


# ---------------------------------------------------------------------------
# Boundary conditions, explicit xy-diffusion
# ---------------------------------------------------------------------------


def apply_bc_all(u, v, w, top_wall_bc_type="dirichlet"):
    """solver.py:70-107 (in place, same assignment order)."""
    u[0, :, :] = u[-1, :, :]
    u[:, 0, :] = u[:, -2, :]
    u[:, -1, :] = u[:, 1, :]
    u[:, :, 0] = -u[:, :, 1]
    if top_wall_bc_type == "neumann":
        u[:, :, -1] = u[:, :, -2]
    else:
        u[:, :, -1] = -u[:, :, -2]
    v[0, :, :] = v[-2, :, :]
    v[-1, :, :] = v[1, :, :]
    v[:, 0, :] = v[:, -1, :]
    v[:, :, 0] = -v[:, :, 1]
    if top_wall_bc_type == "neumann":
        v[:, :, -1] = v[:, :, -2]
    else:
        v[:, :, -1] = -v[:, :, -2]
    w[0, :, :] = w[-2, :, :]
    w[-1, :, :] = w[1, :, :]
    w[:, 0, :] = w[:, -2, :]
    w[:, -1, :] = w[:, 1, :]
    w[:, :, 0] = 0.0
    w[:, :, -1] = 0.0


def diffusion_xy_u(u, nx, ny, nz, dx, dy, nu):
    """operators.py:293-316."""
    diff_u = np.zeros_like(u)
    u_ext_x = np.concatenate([u, u[1:2, :, :]], axis=0)
    d2u_dx2 = (u_ext_x[2 : nx + 2, 1 : ny + 1, 1 : nz + 1]
               - 2 * u_ext_x[1 : nx + 1, 1 : ny + 1, 1 : nz + 1]
               + u_ext_x[0:nx, 1 : ny + 1, 1 : nz + 1]) / dx**2
    d2u_dy2 = (u[1 : nx + 1, 2 : ny + 2, 1 : nz + 1] - 2 * u[1 : nx + 1, 1 : ny + 1, 1 : nz + 1]
               + u[1 : nx + 1, 0:ny, 1 : nz + 1]) / dy**2
    diff_u[1 : nx + 1, 1 : ny + 1, 1 : nz + 1] = nu * (d2u_dx2 + d2u_dy2)
    return diff_u


def diffusion_xy_v(v, nx, ny, nz, dx, dy, nu):
    """operators.py:319-342."""
    diff_v = np.zeros_like(v)
    d2v_dx2 = (v[2 : nx + 2, 1 : ny + 1, 1 : nz + 1] - 2 * v[1 : nx + 1, 1 : ny + 1, 1 : nz + 1]
               + v[0:nx, 1 : ny + 1, 1 : nz + 1]) / dx**2
    v_ext_y = np.concatenate([v, v[:, 1:2, :]], axis=1)
    d2v_dy2 = (v_ext_y[1 : nx + 1, 2 : ny + 2, 1 : nz + 1]
               - 2 * v_ext_y[1 : nx + 1, 1 : ny + 1, 1 : nz + 1]
               + v_ext_y[1 : nx + 1, 0:ny, 1 : nz + 1]) / dy**2
    diff_v[1 : nx + 1, 1 : ny + 1, 1 : nz + 1] = nu * (d2v_dx2 + d2v_dy2)
    return diff_v


def diffusion_xy_w(w, nx, ny, nz, dx, dy, nu):
    """operators.py:345-367."""
    diff_w = np.zeros_like(w)
    d2w_dx2 = (w[2 : nx + 2, 1 : ny + 1, 1:nz] - 2 * w[1 : nx + 1, 1 : ny + 1, 1:nz]
               + w[0:nx, 1 : ny + 1, 1:nz]) / (dx**2)
    d2w_dy2 = (w[1 : nx + 1, 2 : ny + 2, 1:nz] - 2 * w[1 : nx + 1, 1 : ny + 1, 1:nz]
               + w[1 : nx + 1, 0:ny, 1:nz]) / (dy**2)
    diff_w[1 : nx + 1, 1 : ny + 1, 1:nz] = nu * (d2w_dx2 + d2w_dy2)
    return diff_w


# ---------------------------------------------------------------------------
# Tridiagonal solvers and implicit z-diffusion
# ---------------------------------------------------------------------------


def _shift(x, s, down, fill):
    """tridiag.py:23-33."""
    n = x.shape[-1]
    y = np.full_like(x, fill)
    if s < n:
        if down:
            y[..., s:] = x[..., : n - s]
        else:
            y[..., : n - s] = x[..., s:]
    return y


def pcr_solve(a, b, c, d):
    """tridiag.py:36-61 (parallel cyclic reduction, shared 1-D diagonals)."""
    n = d.shape[-1]
    a, b, c, d = a.copy(), b.copy(), c.copy(), d.copy()
    s = 1
    while s < n:
        a_dn = _shift(a, s, True, 0.0)
        b_dn = _shift(b, s, True, 1.0)
        c_dn = _shift(c, s, True, 0.0)
        a_up = _shift(a, s, False, 0.0)
        b_up = _shift(b, s, False, 1.0)
        c_up = _shift(c, s, False, 0.0)
        d_dn = _shift(d, s, True, 0.0)
        d_up = _shift(d, s, False, 0.0)
        alpha = -a / b_dn
        gamma = -c / b_up
        b = b + alpha * c_dn + gamma * a_up
        d = d + alpha * d_dn + gamma * d_up
        a = alpha * a_dn
        c = gamma * c_up
        s *= 2
    return d / b


def solve_implicit_diffusion_u(u, dt, nx, ny, nz, dz_c, dz_f, nu, theta=0.5,
                               top_wall_bc_type="dirichlet"):
    """operators.py:382-473 (also serves v: operators.py:476-550 is identical
    up to the field)."""
    u_new = u.copy()
    alpha = theta * dt * nu
    dz_left = dz_c[0:nz]
    dz_right = dz_c[1 : nz + 1]
    dz_cell = dz_f[0:nz]
    coeff_lower = -alpha / (dz_left * dz_cell)
    coeff_center = 1.0 + alpha * (1.0 / dz_left + 1.0 / dz_right) / dz_cell
    coeff_upper = -alpha / (dz_right * dz_cell)
    a = coeff_lower.copy()
    b = coeff_center.copy()
    c = coeff_upper.copy()
    a[0] = 0.0
    b[0] = coeff_center[0] - coeff_lower[0]
    if top_wall_bc_type == "neumann":
        b[nz - 1] = coeff_center[nz - 1] + coeff_upper[nz - 1]
    else:
        b[nz - 1] = coeff_center[nz - 1] - coeff_upper[nz - 1]
    c[nz - 1] = 0.0
    d = u[1 : nx + 1, 1 : ny + 1, 1 : nz + 1].copy()
    if theta < 1.0:
        beta = (1.0 - theta) * dt * nu
        U = u[1 : nx + 1, 1 : ny + 1, :]
        d2u = ((U[:, :, 2 : nz + 2] - U[:, :, 1 : nz + 1]) / dz_right
               - (U[:, :, 1 : nz + 1] - U[:, :, 0:nz]) / dz_left) / dz_cell
        d2u[:, :, 0] = ((U[:, :, 2] - U[:, :, 1]) / dz_right[0]
                        - 2.0 * U[:, :, 1] / dz_left[0]) / dz_cell[0]
        if top_wall_bc_type == "neumann":
            d2u[:, :, nz - 1] = (-(U[:, :, nz] - U[:, :, nz - 1]) / dz_left[nz - 1]) / dz_cell[nz - 1]
        else:
            d2u[:, :, nz - 1] = (-2.0 * U[:, :, nz] / dz_right[nz - 1]
                                 - (U[:, :, nz] - U[:, :, nz - 1]) / dz_left[nz - 1]) / dz_cell[nz - 1]
        d += beta * d2u
    x = pcr_solve(a, b, c, d.reshape(nx * ny, nz))
    u_new[1 : nx + 1, 1 : ny + 1, 1 : nz + 1] = x.reshape(nx, ny, nz)
    return u_new


def solve_implicit_diffusion_v(v, dt, nx, ny, nz, dz_c, dz_f, nu, theta=0.5,
                               top_wall_bc_type="dirichlet"):
    """operators.py:476-550 (same system as u on the v interior)."""
    return solve_implicit_diffusion_u(v, dt, nx, ny, nz, dz_c, dz_f, nu, theta,
                                      top_wall_bc_type)


def solve_implicit_diffusion_w(w, dt, nx, ny, nz, dz_c, dz_f, nu, theta=0.5):
    """operators.py:553-630."""
    w_new = w.copy()
    alpha = theta * dt * nu
    n = nz - 1
    dz_left = dz_f[0:n]
    dz_right = dz_f[1:nz]
    dz_cv = dz_c[1:nz]
    coeff_lower = -alpha / (dz_left * dz_cv)
    coeff_center = 1.0 + alpha * (1.0 / dz_left + 1.0 / dz_right) / dz_cv
    coeff_upper = -alpha / (dz_right * dz_cv)
    a = coeff_lower.copy()
    b = coeff_center.copy()
    c = coeff_upper.copy()
    a[0] = 0.0
    c[n - 1] = 0.0
    d = w[1 : nx + 1, 1 : ny + 1, 1:nz].copy()
    if theta < 1.0:
        beta = (1.0 - theta) * dt * nu
        W = w[1 : nx + 1, 1 : ny + 1, :]
        d2w = np.empty((nx, ny, n))
        d2w[:, :, 1 : n - 1] = ((W[:, :, 2:n] - W[:, :, 1 : n - 1]) / dz_right[1 : n - 1]
                                - (W[:, :, 1 : n - 1] - W[:, :, 0 : n - 2]) / dz_left[1 : n - 1]) / dz_cv[1 : n - 1]
        d2w[:, :, 0] = ((W[:, :, 1] - W[:, :, 0]) / dz_right[0] - W[:, :, 0] / dz_left[0]) / dz_cv[0]
        d2w[:, :, n - 1] = (-W[:, :, n - 1] / dz_right[n - 1]
                            - (W[:, :, n - 1] - W[:, :, n - 2]) / dz_left[n - 1]) / dz_cv[n - 1]
        d += beta * d2w
    x = pcr_solve(a, b, c, d.reshape(nx * ny, n))
    w_new[1 : nx + 1, 1 : ny + 1, 1:nz] = x.reshape(nx, ny, n)
    return w_new


# ---------------------------------------------------------------------------
# Divergence, FFT-Poisson, projection, bulk
# ---------------------------------------------------------------------------


def compute_divergence(u, v, w, nx, ny, nz, dx, dy, dz_f):
    """utils.py:166-180."""
    du_dx = (u[1 : nx + 1, 1 : ny + 1, 1 : nz + 1] - u[0:nx, 1 : ny + 1, 1 : nz + 1]) / dx
    dv_dy = (v[1 : nx + 1, 1 : ny + 1, 1 : nz + 1] - v[1 : nx + 1, 0:ny, 1 : nz + 1]) / dy
    dw_dz = (w[1 : nx + 1, 1 : ny + 1, 1 : nz + 1] - w[1 : nx + 1, 1 : ny + 1, 0:nz]) / dz_f[0:nz].reshape(1, 1, -1)
    return du_dx + dv_dy + dw_dz


def initialize_fft_solver(nx, ny, nz, dx, dy, dz_c, dz_f, top_wall_bc_type="dirichlet"):
    """projection.py:24-131 (modified wavenumbers, Neumann ends, (0,0) pin)."""
    kx = np.fft.fftfreq(nx, d=dx / (2 * np.pi))
    ky = np.fft.rfftfreq(ny, d=dy / (2 * np.pi))
    kx_mod = (2.0 / dx) * np.sin(kx * dx / 2.0)
    ky_mod = (2.0 / dy) * np.sin(ky * dy / 2.0)
    nkx, nky = len(kx), len(ky)
    kx_grid, ky_grid = np.meshgrid(kx_mod, ky_mod, indexing="ij")
    k_horiz_sq = kx_grid**2 + ky_grid**2
    dz_left = dz_c[:-1]
    dz_right = dz_c[1:]
    coeff_left = 1.0 / (dz_left * dz_f)
    coeff_right = 1.0 / (dz_right * dz_f)
    tri_a = np.zeros((nkx, nky, nz))
    tri_a[:, :, 1:] = coeff_left[1:].reshape(1, 1, -1)
    tri_c = np.zeros((nkx, nky, nz))
    tri_c[:, :, :-1] = coeff_right[:-1].reshape(1, 1, -1)
    coeff_sum = coeff_left + coeff_right
    tri_b = -(coeff_sum.reshape(1, 1, -1) + k_horiz_sq[:, :, None])
    tri_b[:, :, 0] += coeff_left[0]
    tri_b[:, :, -1] += coeff_right[-1]
    tri_b[0, 0, 0] = 1.0
    tri_c[0, 0, 0] = 0.0
    return {"tri_a": tri_a, "tri_b": tri_b, "tri_c": tri_c, "nx": nx, "ny": ny, "nz": nz,
            "top_wall_bc_type": top_wall_bc_type}


def solve_tridiagonal(a, b, c, d):
    """projection.py:199-247 (Thomas over the last axis, per-system diagonals)."""
    batch, n = d.shape
    c_prime = np.zeros((batch, n), dtype=d.dtype)
    d_prime = np.zeros((batch, n), dtype=d.dtype)
    x = np.zeros((batch, n), dtype=d.dtype)
    c_prime[:, 0] = c[:, 0] / b[:, 0]
    d_prime[:, 0] = d[:, 0] / b[:, 0]
    for i in range(1, n):
        denom = b[:, i] - a[:, i] * c_prime[:, i - 1]
        if i < n - 1:
            c_prime[:, i] = c[:, i] / denom
        d_prime[:, i] = (d[:, i] - a[:, i] * d_prime[:, i - 1]) / denom
    x[:, n - 1] = d_prime[:, n - 1]
    for i in range(n - 2, -1, -1):
        x[:, i] = d_prime[:, i] - c_prime[:, i] * x[:, i + 1]
    return x


def solve_poisson_fft(div, fft_data):
    """projection.py:133-197: rfft2 over (x,y), per-mode Thomas, irfft2,
    periodic x/y and Neumann z ghost fill."""
    nx, ny, nz = fft_data["nx"], fft_data["ny"], fft_data["nz"]
    div_hat = np.fft.rfft2(div, axes=(0, 1))
    nkx, nky = div_hat.shape[0], div_hat.shape[1]
    p_hat = solve_tridiagonal(
        fft_data["tri_a"].reshape(-1, nz), fft_data["tri_b"].reshape(-1, nz),
        fft_data["tri_c"].reshape(-1, nz), div_hat.reshape(-1, nz),
    ).reshape(nkx, nky, nz)
    p_interior = np.fft.irfft2(p_hat, s=(nx, ny), axes=(0, 1))
    p = np.zeros((nx + 2, ny + 2, nz + 2))
    p[1 : nx + 1, 1 : ny + 1, 1 : nz + 1] = p_interior
    p[0] = p[nx]
    p[nx + 1] = p[1]
    p[:, 0] = p[:, ny]
    p[:, ny + 1] = p[:, 1]
    p[:, :, 0] = p[:, :, 1]
    p[:, :, nz + 1] = p[:, :, nz]
    return p


def project_velocity(u, v, w, p, nx, ny, nz, dx, dy, dz_c, dz_f, dt):
    """projection.py:250-272 (in place)."""
    dp_dx = (p[2 : nx + 2, 1 : ny + 1, 1 : nz + 1] - p[1 : nx + 1, 1 : ny + 1, 1 : nz + 1]) / dx
    u[1 : nx + 1, 1 : ny + 1, 1 : nz + 1] -= dt * dp_dx
    dp_dy = (p[1 : nx + 1, 2 : ny + 2, 1 : nz + 1] - p[1 : nx + 1, 1 : ny + 1, 1 : nz + 1]) / dy
    v[1 : nx + 1, 1 : ny + 1, 1 : nz + 1] -= dt * dp_dy
    dp_dz = (p[1 : nx + 1, 1 : ny + 1, 2 : nz + 1] - p[1 : nx + 1, 1 : ny + 1, 1:nz]) / dz_c[1:nz].reshape(1, 1, -1)
    w[1 : nx + 1, 1 : ny + 1, 1:nz] -= dt * dp_dz
    return u, v, w


def compute_bulk_velocity(u, cell_vol_ratio, total_volume):
    """utils.py:147-163."""
    nx, ny, nz = cell_vol_ratio.shape
    return np.sum(u[1 : nx + 1, 1 : ny + 1, 1 : nz + 1] * cell_vol_ratio) / total_volume


# ---------------------------------------------------------------------------
# Eulerian IMEX right-hand side
# ---------------------------------------------------------------------------


def compute_momentum_rhs_fused_imex(u, v, w, nx, ny, nz, dx, dy, dz_c, dz_f, nu):
    """operators.py:639-835: divergence-form advection on the MAC grid and the
    xy Laplacian (no z-diffusion), rhs = diffusion_xy - advection on each
    component's interior, zero elsewhere; u and v extended periodically by
    one x-plane / y-plane so every interior face carries its flux."""
    rhs_u, rhs_v, rhs_w = np.zeros_like(u), np.zeros_like(v), np.zeros_like(w)
    dx_inv, dy_inv = 1.0 / dx, 1.0 / dy
    nu_dx2, nu_dy2 = nu * (1.0 / (dx * dx)), nu * (1.0 / (dy * dy))
    dz_f_inv = (1.0 / np.asarray(dz_f)).reshape(1, 1, -1)
    dz_c = np.asarray(dz_c)
    I, J, K = slice(1, nx + 1), slice(1, ny + 1), slice(1, nz + 1)
    # u (operators.py:679-728)
    ue = np.concatenate([u, u[1:2]], axis=0)
    a = 0.5 * (ue[I, J, K] + ue[2:nx + 2, J, K])
    b = 0.5 * (ue[0:nx, J, K] + ue[I, J, K])
    duudx = (a * a - b * b) * dx_inv
    vi = 0.5 * (v[I, J, K] + v[2:nx + 2, J, K])
    ui = 0.5 * (ue[I, J, K] + ue[I, 2:ny + 2, K])
    top = vi * ui
    vi = 0.5 * (v[I, 0:ny, K] + v[2:nx + 2, 0:ny, K])
    ui = 0.5 * (ue[I, 0:ny, K] + ue[I, J, K])
    dvudy = (top - vi * ui) * dy_inv
    wi = 0.5 * (w[I, J, K] + w[2:nx + 2, J, K])
    ui = 0.5 * (ue[I, J, K] + ue[I, J, 2:nz + 2])
    top = wi * ui
    wi = 0.5 * (w[I, J, 0:nz] + w[2:nx + 2, J, 0:nz])
    ui = 0.5 * (ue[I, J, 0:nz] + ue[I, J, K])
    dwudz = (top - wi * ui) * dz_f_inv[:, :, 0:nz]
    adv = duudx + dvudy + dwudz
    d2x = (ue[2:nx + 2, J, K] - 2.0 * ue[I, J, K] + ue[0:nx, J, K]) * nu_dx2
    d2y = (u[I, 2:ny + 2, K] - 2.0 * u[I, J, K] + u[I, 0:ny, K]) * nu_dy2
    rhs_u[I, J, K] = (d2x + d2y) - adv
    # v (operators.py:734-781)
    ve = np.concatenate([v, v[:, 1:2]], axis=1)
    ui = 0.5 * (u[I, J, K] + u[I, 2:ny + 2, K])
    vi = 0.5 * (ve[I, J, K] + ve[2:nx + 2, J, K])
    right = ui * vi
    ui = 0.5 * (u[0:nx, J, K] + u[0:nx, 2:ny + 2, K])
    vi = 0.5 * (ve[0:nx, J, K] + ve[I, J, K])
    duvdx = (right - ui * vi) * dx_inv
    a = 0.5 * (ve[I, J, K] + ve[I, 2:ny + 2, K])
    b = 0.5 * (ve[I, 0:ny, K] + ve[I, J, K])
    dvvdy = (a * a - b * b) * dy_inv
    wi = 0.5 * (w[I, J, K] + w[I, 2:ny + 2, K])
    vi = 0.5 * (ve[I, J, K] + ve[I, J, 2:nz + 2])
    top = wi * vi
    wi = 0.5 * (w[I, J, 0:nz] + w[I, 2:ny + 2, 0:nz])
    vi = 0.5 * (ve[I, J, 0:nz] + ve[I, J, K])
    dwvdz = (top - wi * vi) * dz_f_inv[:, :, 0:nz]
    adv = duvdx + dvvdy + dwvdz
    d2x = (v[2:nx + 2, J, K] - 2.0 * v[I, J, K] + v[0:nx, J, K]) * nu_dx2
    d2y = (ve[I, 2:ny + 2, K] - 2.0 * ve[I, J, K] + ve[I, 0:ny, K]) * nu_dy2
    rhs_v[I, J, K] = (d2x + d2y) - adv
    # w (operators.py:787-833), interior faces 1..nz-1
    Kw = slice(1, nz)
    ui = 0.5 * (u[I, J, Kw] + u[I, J, 2:nz + 1])
    wi = 0.5 * (w[I, J, Kw] + w[2:nx + 2, J, Kw])
    right = ui * wi
    ui = 0.5 * (u[0:nx, J, Kw] + u[0:nx, J, 2:nz + 1])
    wi = 0.5 * (w[0:nx, J, Kw] + w[I, J, Kw])
    duwdx = (right - ui * wi) * dx_inv
    vi = 0.5 * (v[I, J, Kw] + v[I, J, 2:nz + 1])
    wi = 0.5 * (w[I, J, Kw] + w[I, 2:ny + 2, Kw])
    top = vi * wi
    vi = 0.5 * (v[I, 0:ny, Kw] + v[I, 0:ny, 2:nz + 1])
    wi = 0.5 * (w[I, 0:ny, Kw] + w[I, J, Kw])
    dvwdy = (top - vi * wi) * dy_inv
    a = 0.5 * (w[I, J, Kw] + w[I, J, 2:nz + 1])
    b = 0.5 * (w[I, J, 0:nz - 1] + w[I, J, Kw])
    dwwdz = (a * a - b * b) / dz_c[1:nz].reshape(1, 1, -1)
    adv = duwdx + dvwdy + dwwdz
    d2x = (w[2:nx + 2, J, Kw] - 2.0 * w[I, J, Kw] + w[0:nx, J, Kw]) * nu_dx2
    d2y = (w[I, 2:ny + 2, Kw] - 2.0 * w[I, J, Kw] + w[I, 0:ny, Kw]) * nu_dy2
    rhs_w[I, J, Kw] = (d2x + d2y) - adv
    return rhs_u, rhs_v, rhs_w


# ---------------------------------------------------------------------------
# Run-loop diagnostics
# ---------------------------------------------------------------------------


def compute_u_tau(u, z_c, nu, top_wall_bc_type="dirichlet"):
    """utils.py:116-146: plane means of u[:, :, 1] and u[:, :, -2] over the
    whole ghost-shaped x-y extent."""
    dist = z_c[1]
    u_tau_bot = np.sqrt(np.abs(nu * np.mean(u[:, :, 1]) / dist))
    if top_wall_bc_type == "neumann":
        return u_tau_bot
    u_tau_top = np.sqrt(np.abs(nu * np.mean(u[:, :, -2]) / dist))
    return 0.5 * (u_tau_bot + u_tau_top)


def compute_cfl_fused(u, v, w, nx, ny, nz, dx, dy, dz_f, dz_c):
    """operators.py:839-940: max over cells of the x, y and z inverse
    time-scales (4-point averages of the staggered components)."""
    dxi, dyi = 1.0 / dx, 1.0 / dy
    dzfi = (1.0 / np.asarray(dz_f))[0:nz].reshape(1, 1, -1)
    dzci = (1.0 / np.asarray(dz_c))[1:nz].reshape(1, 1, -1)
    I, J, K = slice(1, nx + 1), slice(1, ny + 1), slice(1, nz + 1)
    ux = np.abs(u[I, J, K])
    vx = 0.25 * np.abs(v[I, J, K] + v[I, 0:ny, K] + v[2:nx + 2, J, K] + v[2:nx + 2, 0:ny, K])
    wx = 0.25 * np.abs(w[I, J, K] + w[I, J, 0:nz] + w[2:nx + 2, J, K] + w[2:nx + 2, J, 0:nz])
    dtix = ux * dxi + vx * dyi + wx * dzfi
    uy = 0.25 * np.abs(u[I, J, K] + u[I, 2:ny + 2, K] + u[0:nx, 2:ny + 2, K] + u[0:nx, J, K])
    vy = np.abs(v[I, J, K])
    wy = 0.25 * np.abs(w[I, J, K] + w[I, 2:ny + 2, K] + w[I, 2:ny + 2, 0:nz] + w[I, J, 0:nz])
    dtiy = uy * dxi + vy * dyi + wy * dzfi
    Kz, Kz1 = slice(1, nz), slice(2, nz + 1)
    uz = 0.25 * np.abs(u[I, J, Kz] + u[0:nx, J, Kz] + u[0:nx, J, Kz1] + u[I, J, Kz1])
    vz = 0.25 * np.abs(v[I, J, Kz] + v[I, 0:ny, Kz] + v[I, 0:ny, Kz1] + v[I, J, Kz1])
    wz = np.abs(w[I, J, Kz])
    dtiz = uz * dxi + vz * dyi + wz * dzci
    parts = [dtix.max(), dtiy.max()] + ([dtiz.max()] if nz > 1 else [])
    return float(np.max(parts))


def run_diagnostics(u, v, w, p, nx, ny, nz, dx, dy, dz_f, dz_c, z_c, nu, dt,
                    top_wall_bc_type="dirichlet"):
    """The run loop's per-n_out diagnostics and NaN/Inf guard
    (solver.py:1355-1392, 1436-1479): [max|div|, u_tau, dti*dt, dti,
    non-finite entries of u, v, w, p]."""
    div = compute_divergence(u, v, w, nx, ny, nz, dx, dy, dz_f)
    dti = compute_cfl_fused(u, v, w, nx, ny, nz, dx, dy, dz_f, dz_c)
    bad = sum(int(np.count_nonzero(~np.isfinite(a))) for a in (u, v, w, p) if a is not None)
    return np.array([np.max(np.abs(div)), compute_u_tau(u, z_c, nu, top_wall_bc_type),
                     dti * dt, dti, float(bad)])


# ---------------------------------------------------------------------------
# The step
# ---------------------------------------------------------------------------


def interp1d_weights(coords, nodes):
    """initflow.py:20-28: lower index and upper weight of each coordinate in
    the sorted nodes (searchsorted left, clamped into [1, len-1]; weights
    clamped to [0, 1]: constant extrapolation outside)."""
    k = np.clip(np.searchsorted(nodes, coords, side="left"), 1, len(nodes) - 1)
    k0 = k - 1
    denom = np.maximum(nodes[k] - nodes[k0], 1e-300)
    t = np.clip((coords - nodes[k0]) / denom, 0.0, 1.0)
    return k0, t


def trilinear_tensor(field, xn, yn, zn, xq, yq, zq):
    """initflow.py:31-44: trilinear interpolation of field (len xn, yn, zn)
    onto the tensor-product query nodes, accumulated in (a, b, c) order."""
    ix, tx = interp1d_weights(xq, xn)
    iy, ty = interp1d_weights(yq, yn)
    iz, tz = interp1d_weights(zq, zn)
    out = np.zeros((len(xq), len(yq), len(zq)))
    for a in (0, 1):
        wx = (tx if a else 1.0 - tx).reshape(-1, 1, 1)
        for b in (0, 1):
            wy = (ty if b else 1.0 - ty).reshape(1, -1, 1)
            for c in (0, 1):
                wz = (tz if c else 1.0 - tz).reshape(1, 1, -1)
                out += wx * wy * wz * field[np.ix_(ix + a, iy + b, iz + c)]
    return out


def interpolate_fields(u_s, v_s, w_s, z_c_s, z_f_s, Lx_s, Ly_s, nx, ny, nz, Lx, Ly, z_c, z_f):
    """initflow.py:71-117 (initialize_flow_interpolated after loading): the
    source's periodic ghosts refreshed, x/y mapped proportionally onto the
    source box, z directly; returns u, v, w on the target staggered grid."""
    u_s, v_s, w_s = u_s.copy(), v_s.copy(), w_s.copy()
    nxs, nys = u_s.shape[0] - 1, v_s.shape[1] - 1
    dxs, dys = Lx_s / nxs, Ly_s / nys
    u_s[0, :, :] = u_s[-1, :, :]
    u_s[:, 0, :] = u_s[:, -2, :]
    u_s[:, -1, :] = u_s[:, 1, :]
    v_s[0, :, :] = v_s[-2, :, :]
    v_s[-1, :, :] = v_s[1, :, :]
    v_s[:, 0, :] = v_s[:, -1, :]
    w_s[0, :, :] = w_s[-2, :, :]
    w_s[-1, :, :] = w_s[1, :, :]
    w_s[:, 0, :] = w_s[:, -2, :]
    w_s[:, -1, :] = w_s[:, 1, :]
    x_face_s = np.arange(nxs + 1, dtype=np.float64) * dxs
    x_cent_s = (np.arange(nxs + 2, dtype=np.float64) - 0.5) * dxs
    y_face_s = np.arange(nys + 1, dtype=np.float64) * dys
    y_cent_s = (np.arange(nys + 2, dtype=np.float64) - 0.5) * dys
    dx, dy = Lx / nx, Ly / ny
    rx, ry = Lx_s / Lx, Ly_s / Ly
    x_face_t = np.arange(nx + 1, dtype=np.float64) * dx * rx
    x_cent_t = (np.arange(nx + 2, dtype=np.float64) - 0.5) * dx * rx
    y_face_t = np.arange(ny + 1, dtype=np.float64) * dy * ry
    y_cent_t = (np.arange(ny + 2, dtype=np.float64) - 0.5) * dy * ry
    x_cent_t = np.clip(x_cent_t, x_cent_s[0], x_cent_s[-1])
    y_cent_t = np.clip(y_cent_t, y_cent_s[0], y_cent_s[-1])
    u = trilinear_tensor(u_s, x_face_s, y_cent_s, z_c_s, x_face_t, y_cent_t, z_c)
    v = trilinear_tensor(v_s, x_cent_s, y_face_s, z_c_s, x_cent_t, y_face_t, z_c)
    w = trilinear_tensor(w_s, x_cent_s, y_cent_s, z_f_s, x_cent_t, y_cent_t, z_f)
    return u, v, w


class ChannelOracle:
    """Restates SLChannelFlow's SL-BDF2 step (solver.py:694-1018) on numpy
    arrays.  Fields are ghost-shaped float64 arrays set by the caller."""

    def __init__(self, nx, ny, nz, Lx, Ly, Lz, gamma, Re, U_bulk=1.0, order=6,
                 stretching_type="symmetric", top_wall_bc_type="dirichlet", n_traj_iters=2,
                 z_f=None, z_c=None, tanh_g=None, bdf2_pressure="noninc",
                 bdf2_xy_rhs="extrap", traj_order=2):
        # solver.py:265-281: the two BDF2 options and their checks
        if bdf2_pressure not in ("noninc", "inc"):
            raise ValueError(f"sl.bdf2_pressure must be 'noninc' or 'inc', got {bdf2_pressure}")
        if bdf2_xy_rhs not in ("extrap", "lagged"):
            raise ValueError(f"sl.bdf2_xy_rhs must be 'extrap' or 'lagged', got {bdf2_xy_rhs}")
        self.bdf2_pressure, self.bdf2_xy_rhs = bdf2_pressure, bdf2_xy_rhs
        self._P_curr = self._P_prev = self._p_ext = None
        self.nx, self.ny, self.nz = nx, ny, nz
        self.Lx, self.Ly, self.Lz = Lx, Ly, Lz
        self.nu = 1.0 / Re
        self.U_bulk = U_bulk
        self.top_wall_bc_type = top_wall_bc_type
        if z_f is None:
            z_f, z_c, _, _ = generate_grid(gamma, nz, Lz, stretching_type)
        self.z_f = np.asarray(z_f, dtype=np.float64)
        self.z_c = np.asarray(z_c, dtype=np.float64)
        self.dz_f = self.z_f[1:] - self.z_f[:-1]
        self.dz_c = self.z_c[1:] - self.z_c[:-1]
        self.dx = Lx / nx
        self.dy = Ly / ny
        self.cell_vol_ratio = np.broadcast_to(
            (self.dx * self.dy * self.dz_f).reshape(1, 1, -1), (nx, ny, nz))
        self.total_volume = Lx * Ly * Lz
        self.fft_data = initialize_fft_solver(nx, ny, nz, self.dx, self.dy, self.dz_c,
                                              self.dz_f, top_wall_bc_type)
        self.sl = SLOracle(nx, ny, nz, self.dx, self.dy, Lx, Ly, Lz, self.z_f, self.z_c,
                           gamma, stretching_type=stretching_type, order=order,
                           n_traj_iters=n_traj_iters, traj_order=traj_order)
        if tanh_g is not None:
            self.sl.set_tanh_g(tanh_g)
        self.u = self.v = self.w = self.p = None
        self.u_nm1 = None
        self._Rxy_nm1 = None
        self._dt_prev_step = None
        self.forcing = 0.0
        self.n_clamped_last = 0

    def set_fields(self, u, v, w, p=None):
        self.u, self.v, self.w = (np.array(a, dtype=np.float64) for a in (u, v, w))
        self.p = None if p is None else np.array(p, dtype=np.float64)

    def set_history(self, u_nm1, v_nm1, w_nm1, dt_prev):
        """Adopt V^{n-1} (and the dt it was taken with) so the next step is a
        BDF2 step; R^{n-1} = nu L_xy V^{n-1} exactly as the reference stores
        it (solver.py:847-857 recomputes nothing: the stored R^{n-1} is this)."""
        self.u_nm1 = tuple(np.array(a, dtype=np.float64) for a in (u_nm1, v_nm1, w_nm1))
        a = (self.nx, self.ny, self.nz, self.dx, self.dy, self.nu)
        self._Rxy_nm1 = (diffusion_xy_u(self.u_nm1[0], *a), diffusion_xy_v(self.u_nm1[1], *a),
                         diffusion_xy_w(self.u_nm1[2], *a))
        self._dt_prev_step = dt_prev

    def apply_bc_uvw(self):
        apply_bc_all(self.u, self.v, self.w, self.top_wall_bc_type)

    def _explicit_xy_rhs(self):
        """solver.py:680-688."""
        a = (self.nx, self.ny, self.nz, self.dx, self.dy, self.nu)
        return (diffusion_xy_u(self.u, *a), diffusion_xy_v(self.v, *a),
                diffusion_xy_w(self.w, *a))

    def project_initial(self):
        """solver.py:599-618: project the initial field with dt=1."""
        div = compute_divergence(self.u, self.v, self.w, self.nx, self.ny, self.nz,
                                 self.dx, self.dy, self.dz_f)
        self.p = solve_poisson_fft(div, self.fft_data)
        project_velocity(self.u, self.v, self.w, self.p, self.nx, self.ny, self.nz,
                         self.dx, self.dy, self.dz_c, self.dz_f, 1.0)
        self.apply_bc_uvw()

    def _apply_bulk_forcing(self, dt):
        """solver.py:694-723."""
        nx, ny, nz = self.nx, self.ny, self.nz
        u_bulk_star = compute_bulk_velocity(self.u, self.cell_vol_ratio, self.total_volume)
        correction = self.U_bulk - u_bulk_star
        self.u[1 : nx + 1, 1 : ny + 1, 1 : nz + 1] += correction
        self.apply_bc_uvw()
        self.forcing = correction / dt
        u_bulk_current = compute_bulk_velocity(self.u, self.cell_vol_ratio, self.total_volume)
        return u_bulk_current, self.forcing

    def _zsolve_all(self, dt_z, theta=1.0):
        nx, ny, nz = self.nx, self.ny, self.nz
        self.u = solve_implicit_diffusion_u(self.u, dt_z, nx, ny, nz, self.dz_c, self.dz_f,
                                            self.nu, theta=theta,
                                            top_wall_bc_type=self.top_wall_bc_type)
        self.v = solve_implicit_diffusion_v(self.v, dt_z, nx, ny, nz, self.dz_c, self.dz_f,
                                            self.nu, theta=theta,
                                            top_wall_bc_type=self.top_wall_bc_type)
        self.w = solve_implicit_diffusion_w(self.w, dt_z, nx, ny, nz, self.dz_c, self.dz_f,
                                            self.nu, theta=theta)
        self.apply_bc_uvw()

    def step_imex(self, dt):
        """solver.py:1055-1127: AB2 explicit advection + xy-diffusion (forward
        Euler on the first step), Crank-Nicolson z-diffusion, projection with
        dt, bulk shift."""
        self.apply_bc_uvw()
        rhs = compute_momentum_rhs_fused_imex(self.u, self.v, self.w, self.nx, self.ny, self.nz,
                                              self.dx, self.dy, self.dz_c, self.dz_f, self.nu)
        prev = getattr(self, "_imex_prev", None)
        if prev is None:
            self.u = self.u + dt * rhs[0]
            self.v = self.v + dt * rhs[1]
            self.w = self.w + dt * rhs[2]
        else:
            self.u = self.u + dt * (1.5 * rhs[0] - 0.5 * prev[0])
            self.v = self.v + dt * (1.5 * rhs[1] - 0.5 * prev[1])
            self.w = self.w + dt * (1.5 * rhs[2] - 0.5 * prev[2])
        self._imex_prev = rhs
        self.apply_bc_uvw()
        self._zsolve_all(dt, theta=0.5)
        self._project(dt, dt)
        return self._apply_bulk_forcing(dt)

    def _project(self, rhs_scale, dt_proj):
        nx, ny, nz = self.nx, self.ny, self.nz
        div = compute_divergence(self.u, self.v, self.w, nx, ny, nz, self.dx, self.dy, self.dz_f)
        phi = solve_poisson_fft(div / rhs_scale, self.fft_data)
        project_velocity(self.u, self.v, self.w, phi, nx, ny, nz, self.dx, self.dy,
                         self.dz_c, self.dz_f, dt_proj)
        self.p = phi
        self.apply_bc_uvw()

    def _step_sl_bdf1_bootstrap(self, dt):
        """solver.py:729-811 (U* = V^n, one foot at dt, theta=1 at dt)."""
        nx, ny, nz = self.nx, self.ny, self.nz
        rhs = self._explicit_xy_rhs()
        self.u_nm1 = (self.u.copy(), self.v.copy(), self.w.copy())
        self._Rxy_nm1 = tuple(r.copy() for r in rhs)
        us, vs, ws = self.sl.advect(self.u, self.v, self.w, self.u, self.v, self.w, dt)
        us[1 : nx + 1, 1 : ny + 1, 1 : nz + 1] += dt * rhs[0][1 : nx + 1, 1 : ny + 1, 1 : nz + 1]
        vs[1 : nx + 1, 1 : ny + 1, 1 : nz + 1] += dt * rhs[1][1 : nx + 1, 1 : ny + 1, 1 : nz + 1]
        ws[1 : nx + 1, 1 : ny + 1, 1:nz] += dt * rhs[2][1 : nx + 1, 1 : ny + 1, 1:nz]
        self.n_clamped_last = self.sl.n_clamped_last
        self.u, self.v, self.w = us, vs, ws
        self.apply_bc_uvw()
        self._zsolve_all(dt)
        self._project(dt, dt)
        if self.bdf2_pressure == "inc":  # solver.py:802-807: p^1 = phi, no p^0
            self._P_curr = self.p.copy()
            self._P_prev = None
        self._dt_prev_step = dt
        return self._apply_bulk_forcing(dt)

    def bdf2_predictor(self, dt):
        """solver.py:846-903: the BDF2 predictor (4/3 near - 1/3 far + (2dt/3)
        Rh) of the current state and history, Rh = 2R^n - R^{n-1} ('extrap')
        or R^n ('lagged').  Leaves the state alone; returns (us, vs, ws, R^n)."""
        nx, ny, nz = self.nx, self.ny, self.nz
        rhs = self._explicit_xy_rhs()
        if self.bdf2_xy_rhs == "extrap":
            Rh = tuple(2.0 * r - rn for r, rn in zip(rhs, self._Rxy_nm1))
        else:
            Rh = rhs
        mid = tuple(2.0 * a - b for a, b in zip((self.u, self.v, self.w), self.u_nm1))
        far = self.sl.advect(*self.u_nm1, *mid, 2.0 * dt)
        far = tuple(a.copy() for a in far)
        nc_far = self.sl.n_clamped_last
        near = self.sl.advect(self.u, self.v, self.w, *mid, dt)
        self.n_clamped_last = self.sl.n_clamped_last + nc_far
        dt_eff = (2.0 / 3.0) * dt
        us, vs, ws = (a * (4.0 / 3.0) + (-1.0 / 3.0) * b for a, b in zip(near, far))
        us[1 : nx + 1, 1 : ny + 1, 1 : nz + 1] += dt_eff * Rh[0][1 : nx + 1, 1 : ny + 1, 1 : nz + 1]
        vs[1 : nx + 1, 1 : ny + 1, 1 : nz + 1] += dt_eff * Rh[1][1 : nx + 1, 1 : ny + 1, 1 : nz + 1]
        ws[1 : nx + 1, 1 : ny + 1, 1:nz] += dt_eff * Rh[2][1 : nx + 1, 1 : ny + 1, 1:nz]
        return us, vs, ws, rhs

    def pressure_extrapolation(self):
        """solver.py:906-917: p_ext = 2 p^n - p^{n-1}, or p^n after the
        bootstrap ('inc' only; None before any step)."""
        if self.bdf2_pressure != "inc" or self._P_curr is None:
            return None
        if self._P_prev is None:
            return self._P_curr
        return self._P_curr * 2.0 + (-1.0) * self._P_prev

    def step_sl_bdf2(self, dt):
        """solver.py:813-1018, both BDF2 options (bdf2_pressure 'noninc' /
        'inc', bdf2_xy_rhs 'extrap' / 'lagged')."""
        nx, ny, nz = self.nx, self.ny, self.nz
        self.apply_bc_uvw()
        if (self.u_nm1 is None or self._dt_prev_step is None
                or abs(dt - self._dt_prev_step) > 1e-12 * dt):
            return self._step_sl_bdf1_bootstrap(dt)
        us, vs, ws, rhs = self.bdf2_predictor(dt)
        self._Rxy_nm1 = tuple(r.copy() for r in rhs)
        dt_eff = (2.0 / 3.0) * dt
        p_ext = self.pressure_extrapolation()
        if p_ext is not None:
            # solver.py:918-941: - grad p_ext (project_velocity of zeros with
            # dt = 1), added with dt_eff before the implicit solve
            gp = [np.zeros_like(a) for a in (us, vs, ws)]
            project_velocity(*gp, p_ext, nx, ny, nz, self.dx, self.dy, self.dz_c, self.dz_f,
                             1.0)
            us += dt_eff * gp[0]
            vs += dt_eff * gp[1]
            ws += dt_eff * gp[2]
        self.u_nm1 = (self.u.copy(), self.v.copy(), self.w.copy())
        self.u, self.v, self.w = us, vs, ws
        self.apply_bc_uvw()
        self._zsolve_all(dt_eff)
        self._project(dt_eff, dt_eff)
        if self.bdf2_pressure == "inc":
            # solver.py:999-1012: p^{n+1} = p_ext + phi, history rotated
            phi = self.p
            if self._P_curr is None:
                self._P_curr = phi.copy()
            else:
                self._P_prev = self._P_curr
                self._P_curr = p_ext + phi
            self.p = self._P_curr
        self._dt_prev_step = dt
        return self._apply_bulk_forcing(dt)
