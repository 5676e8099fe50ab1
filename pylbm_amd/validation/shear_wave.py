"""Shear-wave decay: measured vs theoretical viscosity nu = (tau - 1/2)/3.

Same measurement as pylbm/validation/shear_wave.py:11-43: u_x = A sin(k y) in a periodic
n x n box decays as exp(-nu k^2 t); a least-squares fit of the log amplitude of the
x-averaged profile projected on the mode gives nu.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from ..problem import Problem
from . import make_solver


def shear_wave_problem(n: int, tau: float, amplitude: float = 0.01, collision: str = "bgk",
                       **kw: Any) -> Problem:
    k = 2 * np.pi / n
    u0 = np.zeros((2, n, n))
    u0[0] = amplitude * np.sin(k * np.arange(n))[:, None]
    return Problem((n, n), tau=tau, collision=collision, periodic=(True, True), init_u=u0,
                   **kw)


def run_shear_wave(tau: float, n: int = 64, amplitude: float = 0.01, steps: int | None = None,
                   samples: int = 20, collision: str = "bgk", backend: str | None = None,
                   variant: str | None = None, factory: Any = None) -> dict:
    p = shear_wave_problem(n, tau, amplitude, collision)
    k = 2 * np.pi / n
    if steps is None:
        steps = int(2.0 / (p.nu * k * k))  # about two e-folds
    s = make_solver(p, backend, variant, factory)
    chunk = max(s.step_multiple, steps // samples)
    chunk -= chunk % s.step_multiple
    mode = np.sin(k * np.arange(n))
    t, amp = [], []
    for _ in range(samples + 1):
        _, u = s.macroscopic()
        t.append(s.steps_done)
        amp.append(2.0 / n * float(u[0].mean(axis=1) @ mode))
        s.step(chunk)
    slope = np.polyfit(np.asarray(t, float), np.log(np.abs(amp)), 1)[0]
    nu = -slope / (k * k)
    return {"tau": tau, "n": n, "nu_theory": p.nu, "nu_measured": float(nu),
            "rel_error": float(abs(nu - p.nu) / p.nu), "backend": getattr(s, "name", "?"),
            "t": list(map(float, t)), "amplitude": amp}
