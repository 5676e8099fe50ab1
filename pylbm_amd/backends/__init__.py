"""Backend registry: the reference's (backend, variant) lookup (pylbm/backends/__init__.py:16-90)
for the ``hip`` rung.

Same contract as the reference: ``entries()`` lists the pairs in order, ``get()`` imports
the class lazily (KeyError for an unknown pair), ``status()`` reports every entry and never
raises, ``make()`` refuses an unavailable entry with ``Unsupported``.  The table holds
(module, class) pairs; the same ``("hip", variant) -> "hip_backend:Class"`` lines drop into
the reference's ``_ENTRIES``.
"""

from __future__ import annotations

import importlib
import warnings
from typing import Any

from .base import Solver, Unsupported

_HIP = "hip_backend"
_TABLE: dict[tuple[str, str], tuple[str, str]] = {
    ("hip", "fused"): (_HIP, "HipFused"),
    ("hip", "fused-aa"): (_HIP, "HipFusedAA"),
    ("hip", "fused-fp16"): (_HIP, "HipFusedFP16"),
    ("hip", "fused-bf16"): (_HIP, "HipFusedBF16"),
    ("hip", "fused-aa-fp16"): (_HIP, "HipFusedAAFP16"),
    ("hip", "fused-aa-bf16"): (_HIP, "HipFusedAABF16"),
    ("hip", "gather"): (_HIP, "HipGather"),
    ("hip", "fused-f32s"): (_HIP, "HipFusedF32S"),
    ("hip", "fused-2step"): (_HIP, "HipFused2Step"),
    ("hip", "fused-2step-fp16"): (_HIP, "HipFused2StepFP16"),
    ("hip", "fused-2step-bf16"): (_HIP, "HipFused2StepBF16"),
}

COLORS = {"hip": "#e4002b"}


def entries() -> list[tuple[str, str]]:
    return [*_TABLE]


def get(backend: str, variant: str) -> type[Solver]:
    """The solver class of an entry, imported on first use (KeyError if not registered)."""
    where = _TABLE.get((backend, variant))
    if where is None:
        raise KeyError(f"no solver registered as {backend}/{variant}; "
                       f"known: {', '.join('/'.join(e) for e in _TABLE)}")
    module, cls = where
    with warnings.catch_warnings():  # an optional dependency's import noise is not ours
        warnings.simplefilter("ignore")
        return getattr(importlib.import_module(f".{module}", __name__), cls)


def _probe(entry: tuple[str, str]) -> tuple[bool, str]:
    try:
        return get(*entry).available()
    except Exception as e:  # noqa: BLE001 - status() reports, it never raises
        return False, f"{type(e).__name__}: {e}"


def status() -> list[tuple[str, str, bool, str]]:
    """(backend, variant, runs here, why not) for every entry."""
    return [(b, v, *_probe((b, v))) for b, v in _TABLE]


def available() -> list[tuple[str, str]]:
    return [(b, v) for b, v, ok, _ in status() if ok]


def make(backend: str, variant: str, problem: Any, **kw: Any) -> Solver:
    cls = get(backend, variant)
    ok, why = cls.available()
    if not ok:
        raise Unsupported(f"{backend}/{variant} unavailable: {why}")
    return cls(problem, **kw)


__all__ = ["COLORS", "Solver", "Unsupported", "available", "entries", "get", "make", "status"]
