"""Body-force driven Poiseuille channel against the parabola, and its grid convergence.

The measurement of pylbm/validation/poiseuille.py:11-63: walls (half-way bounce-back) at
y = 0 and y = h, periodic x, a uniform Guo force F along x sized for a centre speed u_max;
at steady state u_x(y) = F / (2 nu) y (h - y) at the node centres y = j + 1/2.  The
x-averaged profile is reduced on the device; h numbers per check leave the GPU.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from ..problem import Problem
from . import make_solver
from ._measure import Steady, fields, solver_name


def analytic(h: int, force: float, nu: float) -> np.ndarray:
    """u_x at y = j + 1/2 for walls at y = 0 and y = h (rho = 1)."""
    y = np.arange(h) + 0.5
    return (force / (2.0 * nu)) * y * (h - y)


def run_poiseuille(h: int, tau: float = 0.8, u_max: float = 0.02, collision: str = "bgk",
                   nx: int = 4, backend: str | None = None, variant: str | None = None,
                   factory: Any = None, tol: float = 1e-9, max_steps: int = 2_000_000) -> dict:
    """Step to steady state; relative L2 / Linf errors of the profile against the parabola,
    and the profile itself (y / h at the nodes, u / u_max).

    The check interval is half a viscous time h^2 / (2 nu); the run stops when the largest
    profile change per check falls below ``tol`` u_max, or at the round-off floor (no
    halving of the change for 5 checks once it is below 1e-4).
    """
    nu = (tau - 0.5) / 3.0
    force = 8.0 * nu * u_max / h ** 2
    p = Problem((nx, h), tau=tau, collision=collision, periodic=(True, False),
                force=(force, 0.0))
    s = make_solver(p, backend, variant, factory)
    m = getattr(s, "step_multiple", 1)
    interval = max(50 * m, int(0.5 * h * h / nu))
    interval += interval % m
    monitor = Steady(tol, improve=0.5, patience=5, floor=1e-4)
    prof = np.zeros(h)
    while s.steps_done < max_steps:
        s.step(interval)
        new = fields(s)[1][0].mean(dim=1).cpu().numpy()
        change = float(np.abs(new - prof).max()) / u_max
        prof = new
        if monitor.update(change) is not None:
            break
    ref = analytic(h, force, nu)
    err = prof - ref
    return {"h": h, "tau": tau, "collision": collision, "backend": solver_name(s),
            "steps": s.steps_done,
            "L2": float(np.linalg.norm(err) / np.linalg.norm(ref)),
            "Linf": float(np.abs(err).max() / ref.max()),
            "y": ((np.arange(h) + 0.5) / h).tolist(), "u": (prof / u_max).tolist()}


def convergence(hs: tuple[int, ...] = (8, 16, 32, 64), **kw: Any) -> dict:
    """Runs at several widths and the fitted order of the L2 error."""
    runs = [run_poiseuille(h, **kw) for h in hs]
    errs = np.array([r["L2"] for r in runs])
    order = None
    if (errs > 0).all():
        lh, le = np.log(np.asarray(hs, float)), np.log(errs)
        dh = lh - lh.mean()
        order = float(-(dh * (le - le.mean())).sum() / (dh * dh).sum())
    return {"runs": runs, "order": order}
