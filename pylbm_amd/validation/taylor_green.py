"""D3Q19 Taylor-Green decay (pylbm/validation/taylor_green.py:1-50).

A 2D Taylor-Green vortex u = U(-cos ka sin kb, sin ka cos kb) embedded in a coordinate
plane of a periodic n^3 box has kinetic energy decaying as exp(-4 nu k^2 t).
"""

from __future__ import annotations

from typing import Any

import numpy as np

from ..problem import Problem
from . import make_solver

PLANES = {"xy": (0, 1), "yz": (1, 2), "xz": (0, 2)}


def tgv_problem(n: int, tau: float, amplitude: float, plane: str, **kw: Any) -> Problem:
    a, b = PLANES[plane]
    k = 2 * np.pi / n
    zz, yy, xx = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    coord = {0: xx, 1: yy, 2: zz}
    u = np.zeros((3, n, n, n))
    u[a] = -amplitude * np.cos(k * coord[a]) * np.sin(k * coord[b])
    u[b] = amplitude * np.sin(k * coord[a]) * np.cos(k * coord[b])
    p0 = -0.25 * amplitude ** 2 * (np.cos(2 * k * coord[a]) + np.cos(2 * k * coord[b]))
    return Problem((n, n, n), tau=tau, periodic=(True, True, True), init_u=u,
                   init_rho=1.0 + 3.0 * p0, **kw)


def run_taylor_green(n: int = 32, tau: float = 0.8, amplitude: float = 0.02, plane: str = "xy",
                     samples: int = 10, backend: str | None = None, variant: str | None = None,
                     factory: Any = None, collision: str = "bgk") -> dict:
    p = tgv_problem(n, tau, amplitude, plane, collision=collision)
    k = 2 * np.pi / n
    steps = int(1.0 / (4 * p.nu * k * k))
    s = make_solver(p, backend, variant, factory)
    chunk = max(s.step_multiple, steps // samples)
    chunk -= chunk % s.step_multiple
    t, e = [], []
    for _ in range(samples + 1):
        rho, u = s.macroscopic()
        t.append(s.steps_done)
        e.append(float(0.5 * (rho * (u ** 2).sum(axis=0)).mean()))
        s.step(chunk)
    slope = np.polyfit(np.asarray(t, float), np.log(e), 1)[0]
    nu = -slope / (4 * k * k)
    return {"n": n, "tau": tau, "plane": plane, "backend": getattr(s, "name", "?"),
            "nu_theory": p.nu, "nu_measured": float(nu),
            "rel_error": float(abs(nu - p.nu) / p.nu)}
