"""CPU oracle for the on-the-fly turbulence statistics -- TEST INFRASTRUCTURE ONLY.

A numpy restatement of ``TurbulenceStats.accumulate_statistics`` and
``finalize_statistics`` (slchannel/turbstats.py:156-361 and 362-467), used
only as the checker of the HIP statistics (``slc_stats_*``): by ``tests/``
and ``__graft_entry__.smoke()``.  The product never imports it.

Per snapshot (all float64, interiors of the ghost-shaped fields):

* plane means of the cell-centred u, v, w (U, V, W) and their time sums
  (U_sum, Vm_sum, Wm_sum, Um2_sum, Vm2_sum, Wm2_sum, UmWm_sum);
* u'u' and v'v' at the u and v nodes, about the node plane mean;
* w'w' at the z faces about the face plane mean, then averaged to centres;
* <u'w'> from the cell-centred fluctuations;
* 2-D spectra (rfft2, unnormalised forward, / (nx ny)^2) of u and v node
  fluctuations and the cell-centred w fluctuation at a plane pair
  (bottom, top; w negated at the top), both walls averaged, folded onto
  kx > 0, ky > 0.

Pinned by closed forms in tests/test_stats_oracle.py: the constructed case of
the reference's tests/test_stats_plane_mean.py (time-mean correction, node vs
cell-centre variance, zero uw/ww), per-level variances of separable
patterns, single Fourier modes at known (kx, ky) with amplitude 1/4, Parseval
over the folded half-plane and the top-wall sign of the uw cospectrum.
"""

import numpy as np

PROFILE_NAMES = ("U_sum", "uu_sum", "vv_sum", "ww_sum", "uw_sum", "Vm_sum", "Wm_sum",
                 "Um2_sum", "Vm2_sum", "Wm2_sum", "UmWm_sum")
SPECTRUM_NAMES = ("uu", "vv", "ww", "uw")


def interiors(u, v, w, nx, ny, nz):
    """turbstats.py:168-170: u_int (nx+1, ny, nz), v_int (nx, ny+1, nz),
    w_int (nx, ny, nz+1)."""
    return (u[0:nx + 1, 1:ny + 1, 1:nz + 1], v[1:nx + 1, 0:ny + 1, 1:nz + 1],
            w[1:nx + 1, 1:ny + 1, 0:nz + 1])


def snapshot(u, v, w, nx, ny, nz):
    """The per-snapshot fields of turbstats.py:172-237: dict with U, V, W,
    uu, vv, ww, uw profiles (nz,) and the fluctuation fields the spectra
    use (u_node_f, v_node_f, w_fluct)."""
    u_int, v_int, w_int = interiors(u, v, w, nx, ny, nz)
    ucc = 0.5 * (u_int[:-1] + u_int[1:])                      # 174-175
    U = ucc.mean(axis=(0, 1))
    u_fluct = ucc - U[None, None, :]                          # 180
    vcc = 0.5 * (v_int[:, :-1] + v_int[:, 1:])                # 184-186
    V = vcc.mean(axis=(0, 1))
    wcc = 0.5 * (w_int[:, :, :-1] + w_int[:, :, 1:])          # 189-191
    W = wcc.mean(axis=(0, 1))
    w_fluct = wcc - W[None, None, :]
    u_node = u_int[1:]                                        # 218-223
    v_node = v_int[:, 1:]
    u_node_f = u_node - u_node.mean(axis=(0, 1), keepdims=True)
    v_node_f = v_node - v_node.mean(axis=(0, 1), keepdims=True)
    uu = (u_node_f * u_node_f).mean(axis=(0, 1))
    vv = (v_node_f * v_node_f).mean(axis=(0, 1))
    w_face_f = w_int - w_int.mean(axis=(0, 1), keepdims=True)  # 229-231
    ww_face = (w_face_f * w_face_f).mean(axis=(0, 1))
    ww = 0.5 * (ww_face[:-1] + ww_face[1:])
    uw = (u_fluct * w_fluct).mean(axis=(0, 1))                # 237
    return dict(U=U, V=V, W=W, uu=uu, vv=vv, ww=ww, uw=uw,
                u_node_f=u_node_f, v_node_f=v_node_f, w_fluct=w_fluct)


def profile_increments(s):
    """The 11 profile sums one snapshot adds, in PROFILE_NAMES order
    (turbstats.py:177, 195-200, 239-242)."""
    U, V, W = s["U"], s["V"], s["W"]
    return np.stack([U, s["uu"], s["vv"], s["ww"], s["uw"], V, W, U * U, V * V, W * W, U * W])


def _auto(F, norm):
    return (F.real * F.real + F.imag * F.imag) / norm        # turbstats.py:288-289


def _cross_re(F, G, norm):
    return (F.real * G.real + F.imag * G.imag) / norm        # turbstats.py:291-293


def fold(E, nx, ny):
    """turbstats.py:318-339: E (nx, ny//2+1) -> (nx//2, ny//2) with
    out[a, b] = E[a+1, b+1] + E[nx-1-a, b+1]."""
    nkx, nky = nx // 2, ny // 2
    pos = E[1:nkx + 1, 1:nky + 1]
    neg = E[nx - nkx:nx, 1:nky + 1][::-1]
    return (pos + neg)[:nkx, :nky]


def spectra_increment(s, k_bot, k_top, nx, ny):
    """Folded uu, vv, ww, uw spectra one snapshot adds at the plane pair
    (k_bot, k_top) -- ghost-offset z indices, 1..nz (turbstats.py:256-351)."""
    norm = float(nx * ny) ** 2
    uf, vf, wf = s["u_node_f"], s["v_node_f"], s["w_fluct"]
    fb = [np.fft.rfft2(a[:, :, k_bot - 1]) for a in (uf, vf, wf)]
    ft = [np.fft.rfft2(uf[:, :, k_top - 1]), np.fft.rfft2(vf[:, :, k_top - 1]),
          np.fft.rfft2(-wf[:, :, k_top - 1])]
    out = []
    for c in range(3):
        out.append(fold(0.5 * (_auto(fb[c], norm) + _auto(ft[c], norm)), nx, ny))
    out.append(fold(0.5 * (_cross_re(fb[0], fb[2], norm) + _cross_re(ft[0], ft[2], norm)),
                    nx, ny))
    return np.stack(out)


class StatsOracle:
    """Accumulator with the reference's state: profile sums (11, nz), spectra
    sums per plane pair (4, nx//2, ny//2), sample counters."""

    def __init__(self, nx, ny, nz, planes=()):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.planes = [(int(a), int(b)) for a, b in planes]
        self.prof = np.zeros((11, nz))
        self.spec = np.zeros((len(self.planes), 4, nx // 2, ny // 2))
        self.n_samples = 0

    def accumulate(self, u, v, w):
        s = snapshot(np.asarray(u, np.float64), np.asarray(v, np.float64),
                     np.asarray(w, np.float64), self.nx, self.ny, self.nz)
        self.prof += profile_increments(s)
        for p, (kb, kt) in enumerate(self.planes):
            self.spec[p] += spectra_increment(s, kb, kt, self.nx, self.ny)
        self.n_samples += 1
        return s

    def finalize(self, nu, z_c1, top_wall_bc_type="dirichlet"):
        """turbstats.py:375-412: means, time-mean correction, u_tau."""
        n = self.n_samples
        if n == 0:
            raise RuntimeError("No samples accumulated - cannot finalize statistics")
        d = dict(zip(PROFILE_NAMES, self.prof))
        U = d["U_sum"] / n
        Vm, Wm = d["Vm_sum"] / n, d["Wm_sum"] / n
        uu = d["uu_sum"] / n + d["Um2_sum"] / n - U ** 2
        vv = d["vv_sum"] / n + d["Vm2_sum"] / n - Vm ** 2
        ww = d["ww_sum"] / n + d["Wm2_sum"] / n - Wm ** 2
        uw = d["uw_sum"] / n + d["UmWm_sum"] / n - U * Wm
        if top_wall_bc_type == "neumann":
            dudz = float(U[0] / z_c1)
        else:
            dudz = float(0.5 * (U[0] + U[-1]) / z_c1)
        return dict(U_mean=U, uu_mean=uu, vv_mean=vv, ww_mean=ww, uw_mean=uw,
                    u_tau=float(np.sqrt(nu * dudz)), spectra=self.spec / n)
