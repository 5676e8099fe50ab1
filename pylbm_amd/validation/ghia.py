"""Lid-driven cavity against Ghia, Ghia & Shin (1982) centrelines (pylbm/validation/ghia.py).

``data/ghia1982.json`` is the published table (Tables I and II of the paper) as the
reference stores it; Re = 400, x = 0.9063 of Table II is a known misprint and is excluded.
The steady-state test (RMS change of u between checks, in units of the lid speed) runs on
the device; only the two centrelines of the final field leave the GPU.
"""

from __future__ import annotations

import json
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any

import numpy as np

from ..problem import Problem
from . import make_solver
from ._measure import Steady, fields, solver_name

_DATA = Path(__file__).parent / "data" / "ghia1982.json"
EXCLUDE = {("v", "400", 0.9063)}


def reference(re: int) -> dict[str, np.ndarray]:
    """y, u(y) on x = 1/2 and x, v(x) on y = 1/2 from the table, for Re in {100, 400, 1000}."""
    table = json.loads(_DATA.read_text())
    col = str(int(re))
    vert, horiz = table["u_vertical"], table["v_horizontal"]
    return {"y": np.asarray(vert["y"]), "u": np.asarray(vert[col]),
            "x": np.asarray(horiz["x"]), "v": np.asarray(horiz[col])}


def _middle(a: np.ndarray, axis: int) -> np.ndarray:
    """The line through the middle of the box along ``axis`` (mean of the two central lines
    for an even extent; the central line itself, exactly, for an odd one)."""
    n = a.shape[axis]
    lo, hi = (n - 1) // 2, n // 2
    return 0.5 * (np.take(a, lo, axis=axis) + np.take(a, hi, axis=axis))


def centerlines(ux: Any, uy: Any, u_lid: float) -> dict[str, np.ndarray]:
    """u(y) on x = 1/2 and v(x) on y = 1/2 in units of the lid speed, at the node centres
    (i + 1/2) / n of the half-way walls, closed by the wall values (u = 1 on the lid)."""
    ux, uy = np.asarray(ux), np.asarray(uy)
    ny, nx = ux.shape
    return {"y": np.r_[0.0, (np.arange(ny) + 0.5) / ny, 1.0],
            "u": np.r_[0.0, _middle(ux, 1) / u_lid, 1.0],
            "x": np.r_[0.0, (np.arange(nx) + 0.5) / nx, 1.0],
            "v": np.r_[0.0, _middle(uy, 0) / u_lid, 0.0]}


def errors(prof: dict[str, np.ndarray], re: int) -> dict[str, float]:
    """RMS and max errors at the interior table points (excluded misprints left out)."""
    ref = reference(re)
    out: dict[str, float] = {}
    for comp, coord in (("u", "y"), ("v", "x")):
        pts = ref[coord]
        use = [j for j, q in enumerate(pts)
               if 0.0 < q < 1.0 and (comp, str(int(re)), float(q)) not in EXCLUDE]
        diff = np.interp(pts[use], prof[coord], prof[comp]) - ref[comp][use]
        out[f"{comp}_L2"] = float(np.sqrt((diff ** 2).mean()))
        out[f"{comp}_Linf"] = float(np.abs(diff).max())
    out["L2"] = float(np.hypot(out["u_L2"], out["v_L2"]) / np.sqrt(2.0))
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
    """Step the cavity to steady state and compare its centrelines with the table.

    Converged when the RMS change of u between checks (every max(1000, 4 n) steps) drops
    below ``tol`` u_lid; "plateau" when it stops improving by 10% for 10 checks while below
    50 tol (the round-off floor of 16-bit storage).
    """
    p = Problem.cavity(n, re=re, u_lid=u_lid, collision=collision)
    s = make_solver(p, backend, variant, factory)
    m = getattr(s, "step_multiple", 1)
    check = check_every or max(1000, 4 * n)
    check -= check % m
    monitor = Steady(tol, improve=0.9, patience=10, floor=50 * tol)
    u_prev = fields(s)[1]
    converged: bool | str = False
    while s.steps_done < max_steps:
        s.step(check)
        u = fields(s)[1]
        if not bool(u.isfinite().all()):
            break
        change = float((u - u_prev).pow(2).mean().sqrt()) / u_lid
        u_prev = u
        state = monitor.update(change)
        if state is not None:
            converged = state
            break
    ux, uy = u_prev[0].cpu().numpy(), u_prev[1].cpu().numpy()
    prof = centerlines(ux, uy, u_lid)
    return CavityResult(re, n, solver_name(s), collision, s.steps_done, converged,
                        errors(prof, re), {k: v.tolist() for k, v in prof.items()})
