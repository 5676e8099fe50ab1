"""Physical validation of the hip backend (mirrors pylbm/validation/__init__.py:1-36).

Every case takes a ``backend``/``variant`` pair, or a ready ``solver`` factory (any object
with ``step``, ``macroscopic``, ``steps_done`` and ``step_multiple``), and measures an
analytic quantity: shear-wave and Taylor-Green viscosity, the Poiseuille parabola, and the
Ghia (1982) cavity centrelines.
"""

from __future__ import annotations

from typing import Any, Callable

from ..backends import Solver, get
from ..problem import Problem

#: preference order when a case does not name a variant
GPU_PREFERENCE = [("hip", "fused"), ("hip", "fused-aa")]


def make_solver(problem: Problem, backend: str | None = None, variant: str | None = None,
                factory: Callable[[Problem], Any] | None = None) -> Solver:
    """Instantiate ``factory(problem)``, ``backend/variant``, or the first available hip
    variant that supports ``problem`` (pylbm/validation/__init__.py:20-36)."""
    if factory is not None:
        return factory(problem)
    if backend is not None:
        cls = get(backend, variant or "fused")
        ok, why = cls.available()
        if not ok:
            raise RuntimeError(f"{backend}/{variant} unavailable: {why}")
        return cls(problem)
    for b, v in GPU_PREFERENCE:
        cls = get(b, v)
        if cls.available()[0] and cls.supports(problem)[0]:
            return cls(problem)
    raise RuntimeError("no backend supports this problem")
