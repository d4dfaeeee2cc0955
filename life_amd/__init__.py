"""LIFE fluid + immersed-boundary hot path on MI355X (gfx950).

The compute path is ``liblife_hip.so`` (hand-written HIP kernels behind a C
ABI, see ``include/life_hip.h``); this package is its Python front-end:
:mod:`life_amd.lib` binds the ABI with ctypes, :mod:`life_amd.case` wraps a
case, :mod:`life_amd.config` describes cases the way the reference's
``params.h`` and ``geometry.config`` do.  There is no CPU fallback: if the
shared library is missing, importing :mod:`life_amd.lib` raises.
"""
from . import config  # noqa: F401

__all__ = ["config", "lib", "case", "build"]


def build(verbose=False):
    """Compile liblife_hip.so for gfx950 in place (hipcc; no GPU needed)."""
    import os
    import subprocess
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc")
    jobs = str(min(8, os.cpu_count() or 1))
    subprocess.check_call(["make", "-j" + jobs, "-C", here] + ([] if verbose else ["-s"]))
