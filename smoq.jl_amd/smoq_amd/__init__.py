"""smoq_amd — host-side mirror of SmoQ.jl's state-vector functions over the
MI355X (gfx950) kernels of libsmoq_amd.so.

Function names and argument order follow the reference's utils/cpu modules
(without Julia's `!`); states are device-resident GPUState objects.
"""
from __future__ import annotations

import os
import subprocess

from ._lib import LIB_PATH, SmoqError, lib  # noqa: F401
from .state import (Context, GPUState, default_context, make_initial_psi,  # noqa: F401
                    make_product_ket)

PKG_ROOT = os.path.normpath(os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))


def build(jobs: int = 8) -> str:
    """Compiles libsmoq_amd.so for gfx950 (hipcc cross-compiles; no GPU needed)."""
    subprocess.run(["make", "-s", "-C", PKG_ROOT, f"-j{jobs}"], check=True)
    return LIB_PATH
