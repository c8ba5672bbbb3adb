"""Adaptive FSP solve and R-step space adapters — mirrors ``src/transientcme/sparse/``.

Placeholder module: the adaptive driver is added in a later commit.
"""
from __future__ import annotations

from . import _lib as L


class AdaptiveFspSparse:
    def __init__(self, ode_method=None, space_adapter=None):
        self.ode_method = ode_method
        self.space_adapter = space_adapter


def solve_adaptive(*args, **kwargs):
    raise L.NumCmeHipError(L.NC_ERR_UNSUPPORTED, "adaptive solve is not available yet")
