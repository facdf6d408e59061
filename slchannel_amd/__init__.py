"""slchannel_amd -- MI355X (gfx950) performance layer for slChannel's
semi-Lagrangian BDF2 step and the pressure projection around it.

Host code is Python on PyTorch-ROCm; the work is done by hand-written HIP
kernels in libslc_hip.so behind the C ABI of include/slc.h, with rocFFT for
the x-y transforms.  Public objects mirror the reference's API:

* :class:`SLAdvector` / :class:`HipSL`   (slchannel.semilag / semilag_triton)
* :class:`SLChannelFlow`                  (slchannel.solver, step_sl_bdf2)
* :class:`HipPlan`                        the plan with every step operator
* :class:`TurbulenceStats`                (slchannel.turbstats; include/slc_stats.h)

Like the reference package, importing it makes float64 the default dtype
(slchannel/__init__.py:22-26): the solver assumes it throughout.
"""

import torch

torch.set_default_dtype(torch.float64)

__version__ = "0.1.0"

from .plan import HipPlan  # noqa: E402
from .semilag import HipSL, SLAdvector  # noqa: E402
from .solver import SLChannelFlow  # noqa: E402
from .turbstats import TurbulenceStats  # noqa: E402

__all__ = ["HipPlan", "HipSL", "SLAdvector", "SLChannelFlow", "TurbulenceStats", "__version__"]
