"""Environment flags of the HIP layer, read once at import.

``SLCHANNEL_HIP`` (default ``1``)
    Select the MI355X HIP kernels for the semi-Lagrangian step (the sibling of
    the reference's ``SLCHANNEL_TRITON``, slchannel/env.py:54-59).  With ``0``
    the layer refuses to build a plan instead of silently running something
    else: there is no CPU fallback in this package.
"""

import os


def refresh():
    """Re-read the flags from the environment (tests and benchmarks only)."""
    global USE_HIP
    USE_HIP = os.environ.get("SLCHANNEL_HIP", "1") == "1"


def summary():
    """One-line description of the active performance layer, for banners."""
    return "hip (gfx950)" if USE_HIP else "none (SLCHANNEL_HIP=0)"


refresh()
