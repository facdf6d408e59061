"""CPU-only checks of the native boundary and the host logic (no GPU calls).

* libslc_hip.so loads and exports every function include/slc.h,
  include/slc_ens.h and include/slc_stats.h declare;
* the ctypes table in slchannel_amd/_lib.py binds exactly those symbols;
* the drop-in classes reject bad arguments with the reference's errors
  (slchannel/semilag.py:204-219) before touching a device;
* with no GPU the layer refuses to run (no CPU fallback).
"""

import os
import re
import subprocess

import pytest
import torch

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    out = set()
    for h in ("slc.h", "slc_ens.h", "slc_stats.h"):
        hdr = open(os.path.join(REPO, "include", h)).read()
        out |= set(re.findall(r"^(?:int|int64_t|const char\*)\s+(slc_\w+)\(", hdr, re.M))
    return out


def test_library_exports_every_declared_symbol(check):
    from slchannel_amd import _lib

    lib = _lib.load()
    check("library loads", lib is not None, _lib.LIB_PATH)
    decl = declared_symbols()
    check("headers declare entry points", len(decl) >= 35, f"{len(decl)} declared")
    missing = [s for s in decl if not hasattr(lib, s)]
    check("every declared symbol exported", not missing, f"missing={missing}")
    check("ctypes table == header", set(_lib.EXPORTS) == decl,
          f"diff={sorted(set(_lib.EXPORTS) ^ decl)}")
    nm = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                        text=True).stdout
    exported = set(re.findall(r" T (slc_\w+)$", nm, re.M))
    check("no undeclared slc_ exports", exported <= decl, f"extra={sorted(exported - decl)}")
    check("version", lib.slc_version() >= 100, str(lib.slc_version()))
    check("last_error empty at start", lib.slc_last_error() == b"", "")


def test_null_and_invalid_arguments_return_einval(check):
    """Host-side argument checks of the ABI need no device."""
    import ctypes

    from slchannel_amd import _lib

    lib = _lib.load()
    rc = lib.slc_plan_create(None, 0, None)
    check("plan_create(NULL)", rc == -1, f"rc={rc} msg={lib.slc_last_error()!r}")
    desc = _lib.GridDesc(8, 8, 8, 1.0, 1.0, 2.0, 1.5, 0.9, 0, 5, 2, 2, 0, 1e-3, None, None)
    h = ctypes.c_void_p()
    rc = lib.slc_plan_create(ctypes.byref(desc), 0, ctypes.byref(h))
    check("order 5 rejected", rc == -1 and b"order" in lib.slc_last_error(), f"rc={rc}")
    desc.order = 6
    desc.traj_order = 3
    rc = lib.slc_plan_create(ctypes.byref(desc), 0, ctypes.byref(h))
    check("traj_order 3 rejected", rc == -1 and b"traj_interp_order" in lib.slc_last_error(),
          f"rc={rc} {lib.slc_last_error()!r}")
    # traj_order 4 (tricubic) is a supported option: what stops this
    # descriptor is the missing z tables, checked before any device work
    desc.traj_order = 4
    rc = lib.slc_plan_create(ctypes.byref(desc), 0, ctypes.byref(h))
    check("traj_order 4 accepted; missing z tables rejected",
          rc == -1 and b"z_f / z_c host tables required" in lib.slc_last_error(),
          f"rc={rc} {lib.slc_last_error()!r}")
    rc = lib.slc_departure(None, 0, None, None, None, 0, 0.1, None, None, None, None, None)
    check("departure: null plan rejected", rc == -1, f"rc={rc}")
    rc = lib.slc_apply_bc(None, None, None, None, None)
    check("null plan rejected", rc == -1, f"rc={rc}")


@pytest.mark.parametrize("kw,msg", [
    (dict(order=5), "interp_order"),
    (dict(traj_order=3), "traj_interp_order"),
    (dict(stretching_type="cosine"), "stretching"),
    (dict(interp_dtype="fp16"), "interp_dtype"),
])
def test_advector_validation_matches_reference(check, kw, msg):
    from slchannel_amd import SLAdvector
    from slchannel_amd.grid import generate_grid

    z_f, z_c, _, _ = generate_grid(1.5, 16, 2.0)
    args = dict(order=4)
    args.update(kw)
    with pytest.raises(ValueError) as e:
        SLAdvector(8, 8, 16, 0.5, 0.5, 4.0, 4.0, 2.0, z_f, z_c, 1.5, **args)
    check(f"ValueError mentions {msg}", msg in str(e.value), str(e.value))


def test_no_gpu_means_no_layer(check):
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible; this checks the CPU-only refusal")
    from slchannel_amd import SLAdvector, SLChannelFlow
    from slchannel_amd.grid import generate_grid

    z_f, z_c, _, _ = generate_grid(1.5, 16, 2.0)
    with pytest.raises(RuntimeError):
        SLAdvector(8, 8, 16, 0.5, 0.5, 4.0, 4.0, 2.0, z_f, z_c, 1.5, order=4)
    cfg = {"grid": {"nx": 8, "ny": 8, "nz": 16}, "domain": {"Lx": 4.0, "Ly": 4.0, "Lz": 2.0},
           "flow": {"Re": 100.0, "gamma": 1.5}}
    with pytest.raises(RuntimeError):
        SLChannelFlow(config=cfg)
    check("HIP layer refuses to run without a GPU", True, "RuntimeError raised")


def test_solver_config_validation(check):
    from slchannel_amd import SLChannelFlow

    base = {"grid": {"nx": 8, "ny": 8, "nz": 16}, "domain": {"Lx": 4.0, "Ly": 4.0, "Lz": 2.0},
            "flow": {"Re": 100.0, "gamma": 1.5}}
    for key, bad, sect in (("stretching_type", "cosine", "domain"),):
        cfg = {k: dict(v) for k, v in base.items()}
        cfg[sect][key] = bad
        with pytest.raises(ValueError):
            SLChannelFlow(config=cfg)
    cfg = {k: dict(v) for k, v in base.items()}
    cfg["boundary_conditions"] = {"top_wall": {"type": "robin"}}
    with pytest.raises(ValueError):
        SLChannelFlow(config=cfg)
    check("invalid configs rejected before any device work", True, "")


def test_grid_matches_oracle_and_reference_formula(check, golden):
    import numpy as np

    from oracle import oracle as orc
    from slchannel_amd.grid import generate_grid

    g = golden("grid_sym_g15_n64")
    z_f, z_c, dz_f, dz_c = generate_grid(1.5, 64, 2.0)
    check("z_f vs reference", np.abs(z_f.numpy() - g["z_f"]).max() < 1e-15, "")
    check("z_c vs reference", np.abs(z_c.numpy() - g["z_c"]).max() < 1e-15, "")
    o = orc.generate_grid(1.5, 64, 2.0)
    check("z_f vs oracle", np.abs(z_f.numpy() - o[0]).max() < 1e-15, "")
