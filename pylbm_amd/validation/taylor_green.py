"""D3Q19 Taylor-Green decay (the measurement of pylbm/validation/taylor_green.py:1-50).

A 2D Taylor-Green vortex u = U(-cos(k a) sin(k b), sin(k a) cos(k b)) placed in one
coordinate plane (a, b) of a periodic n^3 box keeps its shape and loses kinetic energy as
exp(-4 nu k^2 t); the fitted rate over one e-fold gives nu.  The energy is the float64
device reduction of the solver (``reduce("kinetic")``) when it has one.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from ..problem import Problem
from . import make_solver
from ._measure import exp_rate, mean_energy, sample, solver_name

PLANES = {"xy": (0, 1), "yz": (1, 2), "xz": (0, 2)}


def tgv_problem(n: int, tau: float, amplitude: float, plane: str, **kw: Any) -> Problem:
    """The vortex in ``plane`` with its pressure, rho = 1 + 3 p (c_s^2 = 1/3)."""
    a, b = PLANES[plane]
    phase = 2 * np.pi * np.arange(n) / n
    # array axes are (z, y, x): spatial axis s runs along array axis 2 - s
    xa = phase.reshape([n if ax == 2 - a else 1 for ax in range(3)])
    xb = phase.reshape([n if ax == 2 - b else 1 for ax in range(3)])
    u = np.zeros((3, n, n, n))
    u[a] = -amplitude * np.cos(xa) * np.sin(xb)
    u[b] = amplitude * np.sin(xa) * np.cos(xb)
    p0 = -0.25 * amplitude ** 2 * (np.cos(2 * xa) + np.cos(2 * xb))
    rho = 1.0 + 3.0 * np.broadcast_to(p0, (n, n, n))
    return Problem((n, n, n), tau=tau, periodic=(True, True, True), init_u=u,
                   init_rho=np.ascontiguousarray(rho), **kw)


def run_taylor_green(n: int = 32, tau: float = 0.8, amplitude: float = 0.02, plane: str = "xy",
                     samples: int = 10, backend: str | None = None, variant: str | None = None,
                     factory: Any = None, collision: str = "bgk") -> dict:
    p = tgv_problem(n, tau, amplitude, plane, collision=collision)
    k2 = (2 * np.pi / n) ** 2
    horizon = int(1.0 / (4 * p.nu * k2))  # one e-fold of the energy
    s = make_solver(p, backend, variant, factory)
    t, e = sample(s, horizon // samples, samples, lambda: mean_energy(s))
    nu = -exp_rate(t, e) / (4 * k2)
    return {"n": n, "tau": tau, "plane": plane, "backend": solver_name(s),
            "nu_theory": p.nu, "nu_measured": nu, "rel_error": abs(nu - p.nu) / p.nu}
