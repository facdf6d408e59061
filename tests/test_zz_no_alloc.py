"""slc_bulk_force_noalloc: the bulk-flux shift with its partial sums in the
library's static device array instead of the plan's workspace.

Called through the C ABI (ctypes) beside slc_bulk_force on copies of the same
fields: the shifted u (interior and ghosts), v, w and the four output doubles
must be bit-identical.  Two calls in a row, the second on a smaller grid, so
the second call reuses the array the first one left filled.  (No grid needs
more partial sums than the array holds, so the oversize rejection cannot be
reached from a valid plan.)
"""

import ctypes
import math

import pytest
import torch

pytestmark = pytest.mark.gpu


def _plan_and_fields(nx, ny, nz, seed):
    from slchannel_amd.grid import generate_grid
    from slchannel_amd.plan import HipPlan, field_shapes

    z_f, z_c, _, _ = generate_grid(1.5, nz, 2.0)
    plan = HipPlan(nx, ny, nz, 4 * math.pi, 2 * math.pi, 2.0, z_f, z_c, 1.5, order=6,
                   nu=1.0 / 2800.0, device="cuda:0")
    gen = torch.Generator().manual_seed(seed)
    sh = field_shapes(nx, ny, nz)
    u, v, w = (torch.rand(sh[c], generator=gen, dtype=torch.float64).cuda() for c in "uvw")
    return plan, (u, v, w)


def _call(plan, name, fields, U_bulk, dt):
    from slchannel_amd import _lib

    u, v, w = (t.clone() for t in fields)
    out = torch.full((4,), float("nan"), dtype=torch.float64, device="cuda:0")
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = getattr(plan._lib, name)(plan._h, ptr(u), ptr(v), ptr(w), ctypes.c_double(U_bulk),
                                  ctypes.c_double(dt), ptr(out), stream)
    _lib.check(rc, name)
    torch.cuda.synchronize()
    return u, v, w, out


def test_noalloc_bulk_force_matches_plan_workspace(check):
    for nx, ny, nz, seed in ((96, 64, 80, 1), (24, 16, 20, 2)):
        print(f"bulk force no-alloc: grid {nx}x{ny}x{nz} seed {seed}", flush=True)
        plan, fields = _plan_and_fields(nx, ny, nz, seed)
        ref = _call(plan, "slc_bulk_force", fields, 1.0, 0.0125)
        got = _call(plan, "slc_bulk_force_noalloc", fields, 1.0, 0.0125)
        for name, a, b in zip(("u", "v", "w", "out"), got, ref):
            check(f"{nx}x{ny}x{nz}: {name} bit-identical to slc_bulk_force", torch.equal(a, b),
                  f"max|d|={(a - b).abs().max().item():.1e}")
        ub_after = got[3][2].item()
        check(f"{nx}x{ny}x{nz}: bulk velocity 1 after the shift", abs(ub_after - 1.0) < 1e-14,
              f"{ub_after!r}")
        check(f"{nx}x{ny}x{nz}: forcing = correction / dt",
              got[3][1].item() == got[3][0].item() / 0.0125, f"{got[3].tolist()}")
        plan.close()

