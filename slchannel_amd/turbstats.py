"""TurbulenceStats with the per-snapshot work on the GPU (include/slc_stats.h).

Drop-in for slchannel.turbstats.TurbulenceStats (turbstats.py:21-611): same
constructor, same ``accumulate_statistics`` / ``finalize_statistics`` /
``save_statistics`` / ``load_statistics`` / ``save_state`` / ``load_state``
and the same accumulator attributes (``U_sum`` .. ``UmWm_sum``,
``E_uu_2d_sum`` .. ``E_uw_2d_sum``, ``extra_planes``), so state files are
interchangeable with the reference's.  One ``slc_stats_accumulate`` call per
snapshot does the plane reductions and the spectra on the device; the
running sums stay device tensors; ``finalize_statistics`` is the reference's
host arithmetic on nz numbers.

HIP only: the fields must be float64 CUDA tensors; there is no CPU path.
"""

import ctypes
import os

import numpy as np
import torch

from . import _lib

PROFILE_NAMES = ("U_sum", "uu_sum", "vv_sum", "ww_sum", "uw_sum", "Vm_sum", "Wm_sum",
                 "Um2_sum", "Vm2_sum", "Wm2_sum", "UmWm_sum")
SPECTRA = ("uu", "vv", "ww", "uw")


def _as_tensor(z):
    return z if isinstance(z, torch.Tensor) else torch.as_tensor(np.asarray(z), dtype=torch.float64)


class TurbulenceStats:
    """On-the-fly turbulence statistics (turbstats.py:21-154)."""

    def __init__(self, nx, ny, nz, Lx, Ly, Lz, z_c, z_f, dz_c, dz_f, dx, dy, nu, Re_tau_target,
                 z_plus_target=15.0, device="cuda", top_wall_bc_type="dirichlet",
                 spectra_z_planes=None):
        if not torch.cuda.is_available():
            raise RuntimeError("the HIP layer needs a ROCm GPU (torch.cuda.is_available() is False)")
        dev = torch.device(device)
        if dev.type != "cuda":
            raise RuntimeError(f"TurbulenceStats runs on the GPU only, got device={device!r}")
        if dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self._lib = _lib.load()
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.Lx, self.Ly, self.Lz = Lx, Ly, Lz
        self.z_c, self.z_f = _as_tensor(z_c), _as_tensor(z_f)
        self.dz_c, self.dz_f = dz_c, dz_f
        self.dx, self.dy = dx, dy
        self.nu = nu
        self.device = dev
        self.top_wall_bc_type = top_wall_bc_type

        # z+ planes of the spectra (turbstats.py:67-91): closest interior
        # centre to z+ from each wall, ghost-offset index
        delta = Lz / 2.0
        self.u_tau_target = Re_tau_target * nu / delta
        zci = self.z_c[1:self.nz + 1].detach().cpu().to(torch.float64)
        z_t = z_plus_target * nu / self.u_tau_target
        self.k_bot = int(torch.argmin(torch.abs(zci - z_t)).item()) + 1
        self.k_top = int(torch.argmin(torch.abs(zci - (Lz - z_t))).item()) + 1
        self.spectra_z_planes = [float(zp) for zp in (spectra_z_planes or [])]
        extra = []
        for zp in self.spectra_z_planes:
            zt = zp * nu / self.u_tau_target
            kb = int(torch.argmin(torch.abs(zci - zt)).item()) + 1
            kt = int(torch.argmin(torch.abs(zci - (Lz - zt))).item()) + 1
            extra.append((zp, kb, kt))

        self.n_samples = 0
        self.n_samples_extra = 0
        nkx, nky = self.nx // 2, self.ny // 2
        npl = 1 + len(extra)
        f64 = dict(dtype=torch.float64, device=dev)
        self._prof = torch.zeros(len(PROFILE_NAMES), self.nz, **f64)
        self._spec = torch.zeros(npl, 4, nkx, nky, **f64)
        self.extra_planes = [{"zp": zp, "k_bot": kb, "k_top": kt,
                              "sums": {c: self._spec[1 + n, q] for q, c in enumerate(SPECTRA)}}
                             for n, (zp, kb, kt) in enumerate(extra)]
        kb_arr = (ctypes.c_int * npl)(self.k_bot, *[kb for _, kb, _ in extra])
        kt_arr = (ctypes.c_int * npl)(self.k_top, *[kt for _, _, kt in extra])
        h = ctypes.c_void_p()
        with torch.cuda.device(dev):
            _lib.check(self._lib.slc_stats_create(dev.index, self.nx, self.ny, self.nz, npl,
                                                  kb_arr, kt_arr, ctypes.byref(h)),
                       "slc_stats_create")
        self._h = h
        self.kx = 2 * np.pi * np.fft.rfftfreq(self.nx, d=dx)[1:]
        self.ky = 2 * np.pi * np.fft.rfftfreq(self.ny, d=dy)[1:]

    # accumulators as named views of the device sums (turbstats.py:98-128)
    def __getattr__(self, name):
        d = self.__dict__
        if name in PROFILE_NAMES and "_prof" in d:
            return d["_prof"][PROFILE_NAMES.index(name)]
        if (name.startswith("E_") and name.endswith("_2d_sum") and name[2:4] in SPECTRA
                and "_spec" in d):
            return d["_spec"][0, SPECTRA.index(name[2:4])]
        raise AttributeError(name)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.slc_stats_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _field(self, name, t, shape):
        if not isinstance(t, torch.Tensor) or t.device != self.device or t.dtype != torch.float64:
            raise ValueError(f"{name}: float64 tensor on {self.device} required")
        if tuple(t.shape) != shape or not t.is_contiguous():
            raise ValueError(f"{name}: contiguous shape {shape} required, got {tuple(t.shape)}")
        return ctypes.c_void_p(t.data_ptr())

    def accumulate_statistics(self, u, v, w, u_tau_current=None):
        """One snapshot (turbstats.py:156-361); u_tau_current is unused, as
        in the reference (diagnostics only)."""
        nx, ny, nz = self.nx, self.ny, self.nz
        pu = self._field("u", u, (nx + 1, ny + 2, nz + 2))
        pv = self._field("v", v, (nx + 2, ny + 1, nz + 2))
        pw = self._field("w", w, (nx + 2, ny + 2, nz + 1))
        stream = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        _lib.check(self._lib.slc_stats_accumulate(
            self._h, pu, pv, pw, ctypes.c_void_p(self._prof.data_ptr()),
            ctypes.c_void_p(self._spec.data_ptr()), stream), "slc_stats_accumulate")
        if self.extra_planes:
            self.n_samples_extra += 1
        self.n_samples += 1

    def finalize_statistics(self):
        """turbstats.py:362-467."""
        if self.n_samples == 0:
            raise RuntimeError("No samples accumulated - cannot finalize statistics")
        n = self.n_samples
        U = self.U_sum / n
        Vm, Wm = self.Vm_sum / n, self.Wm_sum / n
        uu = self.uu_sum / n + self.Um2_sum / n - U ** 2
        vv = self.vv_sum / n + self.Vm2_sum / n - Vm ** 2
        ww = self.ww_sum / n + self.Wm2_sum / n - Wm ** 2
        uw = self.uw_sum / n + self.UmWm_sum / n - U * Wm
        dz_utau = float(self.z_c[1])
        if self.top_wall_bc_type == "neumann":
            dUdz_wall = float(U[0] / dz_utau)
        else:
            dUdz_wall = float(0.5 * (U[0] + U[-1]) / dz_utau)
        u_tau = float(np.sqrt(self.nu * dUdz_wall))

        def host(t):
            return np.asarray(t.detach().cpu().numpy())

        stats = {
            "n_samples": self.n_samples,
            "z_c": host(self.z_c[1:self.nz + 1]),
            "U_mean": host(U), "uu_mean": host(uu), "vv_mean": host(vv),
            "ww_mean": host(ww), "uw_mean": host(uw),
            "kx": self.kx[:self.nx // 2], "ky": self.ky[:self.ny // 2],
            "E_uu_2d": host(self.E_uu_2d_sum / n), "E_vv_2d": host(self.E_vv_2d_sum / n),
            "E_ww_2d": host(self.E_ww_2d_sum / n), "E_uw_2d": host(self.E_uw_2d_sum / n),
            "z_plus_target": self.k_bot,
            "Re_tau_target": self.u_tau_target * (self.Lz / 2) / self.nu,
            "nu": self.nu, "Re": 1.0 / self.nu, "u_tau": u_tau,
        }
        if self.extra_planes and self.n_samples_extra > 0:
            stats["n_samples_extra"] = self.n_samples_extra
            stats["spectra_z_planes"] = np.asarray(self.spectra_z_planes)
            for pl in self.extra_planes:
                stats[f"z_index_z{pl['zp']:.0f}"] = pl["k_bot"]
                for c in SPECTRA:
                    stats[f"E_{c}_2d_z{pl['zp']:.0f}"] = host(pl["sums"][c] / self.n_samples_extra)
        return stats

    def save_statistics(self, filepath):
        np.savez_compressed(filepath, **self.finalize_statistics())

    @staticmethod
    def load_statistics(filepath):
        if not os.path.exists(filepath):
            raise FileNotFoundError(f"Statistics file not found: {filepath}")
        data = np.load(filepath)
        return {key: data[key] for key in data.files}

    def save_state(self, filepath):
        """Running sums + counters, the reference's state layout
        (turbstats.py:493-535)."""
        state = {"n_samples": self.n_samples}
        for name in PROFILE_NAMES:
            state[name] = np.asarray(getattr(self, name).cpu().numpy())
        for c in SPECTRA:
            state[f"E_{c}_2d_sum"] = np.asarray(getattr(self, f"E_{c}_2d_sum").cpu().numpy())
        state.update(nx=self.nx, ny=self.ny, nz=self.nz)
        if self.extra_planes:
            state["n_samples_extra"] = self.n_samples_extra
            state["spectra_z_planes"] = np.asarray(self.spectra_z_planes)
            for pl in self.extra_planes:
                for c in SPECTRA:
                    state[f"E_{c}_2d_z{pl['zp']:.0f}_sum"] = np.asarray(pl["sums"][c].cpu().numpy())
        np.savez_compressed(filepath, **state)

    def load_state(self, filepath):
        """turbstats.py:536-611: restores sums and counters; the plane-mean
        sums only if the state has them, the extra planes only if the state
        carries the same plane set."""
        if not os.path.exists(filepath):
            raise FileNotFoundError(f"Statistics state file not found: {filepath}")
        data = np.load(filepath)
        if data["nx"] != self.nx or data["ny"] != self.ny or data["nz"] != self.nz:
            raise ValueError(
                f"Grid dimension mismatch! State file has (nx={data['nx']}, ny={data['ny']}, "
                f"nz={data['nz']}) but current simulation has (nx={self.nx}, ny={self.ny}, "
                f"nz={self.nz})")
        self.n_samples = int(data["n_samples"])

        def put(dst, key):
            dst.copy_(torch.as_tensor(data[key], dtype=torch.float64))

        for name in ("U_sum", "uu_sum", "vv_sum", "ww_sum", "uw_sum"):
            put(getattr(self, name), name)
        for c in SPECTRA:
            put(getattr(self, f"E_{c}_2d_sum"), f"E_{c}_2d_sum")
        saved = (list(np.asarray(data["spectra_z_planes"]).ravel())
                 if "spectra_z_planes" in data.files else [])
        if self.extra_planes and saved == self.spectra_z_planes:
            self.n_samples_extra = int(data["n_samples_extra"])
            for pl in self.extra_planes:
                for c in SPECTRA:
                    put(pl["sums"][c], f"E_{c}_2d_z{pl['zp']:.0f}_sum")
        if "Um2_sum" in data:
            for name in ("Vm_sum", "Wm_sum", "Um2_sum", "Vm2_sum", "Wm2_sum", "UmWm_sum"):
                put(getattr(self, name), name)
        else:
            print("  WARNING: state file predates the plane-mean-variance fix: normal stresses "
                  "from these samples are biased low by var_t(plane mean)", flush=True)
