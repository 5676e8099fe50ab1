"""Backend registry (mirrors pylbm/backends/__init__.py:16-90).

Entries are imported lazily; a missing library or GPU only disables the affected entries
(:func:`status` never raises).  This package carries the ``hip`` rung only; the same
``("hip", variant) -> "hip_backend:Class"`` lines drop into the reference's ``_ENTRIES``.
"""

from __future__ import annotations

import importlib
import warnings
from typing import Any

from .base import Solver, Unsupported

_ENTRIES: dict[tuple[str, str], str] = {
    ("hip", "fused"): "hip_backend:HipFused",
    ("hip", "fused-aa"): "hip_backend:HipFusedAA",
    ("hip", "fused-fp16"): "hip_backend:HipFusedFP16",
    ("hip", "fused-bf16"): "hip_backend:HipFusedBF16",
    ("hip", "fused-aa-fp16"): "hip_backend:HipFusedAAFP16",
    ("hip", "fused-aa-bf16"): "hip_backend:HipFusedAABF16",
    ("hip", "gather"): "hip_backend:HipGather",
}

COLORS = {"hip": "#e4002b"}


def entries() -> list[tuple[str, str]]:
    return list(_ENTRIES)


def get(backend: str, variant: str) -> type[Solver]:
    """Import and return a solver class (KeyError for an unknown pair)."""
    try:
        target = _ENTRIES[(backend, variant)]
    except KeyError as e:
        raise KeyError(f"unknown backend/variant {backend}/{variant}") from e
    mod_name, cls_name = target.split(":")
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        mod = importlib.import_module(f"{__name__}.{mod_name}")
    return getattr(mod, cls_name)


def status() -> list[tuple[str, str, bool, str]]:
    out = []
    for b, v in _ENTRIES:
        try:
            ok, why = get(b, v).available()
        except Exception as e:  # noqa: BLE001 - the registry never crashes
            ok, why = False, f"{type(e).__name__}: {e}"
        out.append((b, v, ok, why))
    return out


def available() -> list[tuple[str, str]]:
    return [(b, v) for b, v, ok, _ in status() if ok]


def make(backend: str, variant: str, problem: Any, **kw: Any) -> Solver:
    cls = get(backend, variant)
    ok, why = cls.available()
    if not ok:
        raise Unsupported(f"{backend}/{variant} unavailable: {why}")
    return cls(problem, **kw)


__all__ = ["COLORS", "Solver", "Unsupported", "available", "entries", "get", "make", "status"]
