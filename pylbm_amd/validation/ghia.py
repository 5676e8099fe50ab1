"""Lid-driven cavity vs Ghia, Ghia & Shin (1982) centrelines (pylbm/validation/ghia.py).

``data/ghia1982.json`` is the published table (Tables I and II of the paper) as the
reference stores it; Re = 400, x = 0.9063 of Table II is a known misprint and is excluded.
"""

from __future__ import annotations

import json
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any

import numpy as np

from ..problem import Problem
from . import make_solver

_DATA = Path(__file__).parent / "data" / "ghia1982.json"
EXCLUDE = {("v", "400", 0.9063)}


def reference(re: int) -> dict[str, np.ndarray]:
    t = json.loads(_DATA.read_text())
    key = str(int(re))
    return {"y": np.array(t["u_vertical"]["y"]), "u": np.array(t["u_vertical"][key]),
            "x": np.array(t["v_horizontal"]["x"]), "v": np.array(t["v_horizontal"][key])}


def centerlines(ux: np.ndarray, uy: np.ndarray, u_lid: float) -> dict[str, np.ndarray]:
    """u on the vertical centreline, v on the horizontal one, normalised, with wall values."""
    ny, nx = ux.shape

    def mid(a: np.ndarray, axis: int) -> np.ndarray:
        m = a.shape[axis]
        if m % 2:
            return np.take(a, m // 2, axis=axis)
        return 0.5 * (np.take(a, m // 2 - 1, axis=axis) + np.take(a, m // 2, axis=axis))

    yc = (np.arange(ny) + 0.5) / ny
    xc = (np.arange(nx) + 0.5) / nx
    return {"y": np.concatenate([[0.0], yc, [1.0]]),
            "u": np.concatenate([[0.0], mid(ux, 1) / u_lid, [1.0]]),
            "x": np.concatenate([[0.0], xc, [1.0]]),
            "v": np.concatenate([[0.0], mid(uy, 0) / u_lid, [0.0]])}


def errors(prof: dict[str, np.ndarray], re: int) -> dict[str, float]:
    ref = reference(re)
    out: dict[str, float] = {}
    for comp, coord in (("u", "y"), ("v", "x")):
        pts, vals = ref[coord], ref[comp]
        keep = (pts > 0) & (pts < 1)
        keep &= np.array([(comp, str(re), float(p)) not in EXCLUDE for p in pts])
        d = np.interp(pts[keep], prof[coord], prof[comp]) - vals[keep]
        out[f"{comp}_L2"] = float(np.sqrt(np.mean(d ** 2)))
        out[f"{comp}_Linf"] = float(np.max(np.abs(d)))
    out["L2"] = float(np.sqrt(0.5 * (out["u_L2"] ** 2 + out["v_L2"] ** 2)))
    out["Linf"] = max(out["u_Linf"], out["v_Linf"])
    return out


@dataclass
class CavityResult:
    re: int
    n: int
    backend: str
    collision: str
    steps: int
    converged: bool | str
    errors: dict[str, float]
    profile: dict[str, list[float]] = field(repr=False)


def run_cavity(re: int, n: int = 128, u_lid: float = 0.1, collision: str = "bgk",
               backend: str | None = None, variant: str | None = None, factory: Any = None,
               tol: float = 1e-5, max_steps: int = 1_000_000,
               check_every: int | None = None) -> CavityResult:
    """Run to steady state (RMS change of u per check / U below tol) and compare."""
    p = Problem.cavity(n, re=re, u_lid=u_lid, collision=collision)
    s = make_solver(p, backend, variant, factory)
    check = check_every or max(1000, 4 * n)
    check -= check % s.step_multiple
    _, u_old = s.macroscopic()
    converged: bool | str = False
    best, stale = np.inf, 0
    while s.steps_done < max_steps:
        s.step(check)
        _, u = s.macroscopic()
        if not np.isfinite(u).all():
            break
        change = float(np.sqrt(np.mean((u - u_old) ** 2))) / u_lid
        u_old = u
        if change < tol:
            converged = True
            break
        if change < 0.9 * best:
            best, stale = change, 0
        else:
            stale += 1
        if stale >= 10 and best < 50 * tol:
            converged = "plateau"
            break
    prof = centerlines(u_old[0], u_old[1], u_lid)
    return CavityResult(re, n, getattr(s, "name", "?"), collision, s.steps_done, converged,
                        errors(prof, re), {k: v.tolist() for k, v in prof.items()})
