"""Shear-wave decay: measured against theoretical viscosity nu = (tau - 1/2) / 3.

The measurement of pylbm/validation/shear_wave.py:11-43: u_x = A sin(k y) in a periodic
n x n box decays as exp(-nu k^2 t); the amplitude of the x-averaged u_x on the sine mode is
sampled over about two e-folds and its exponential rate gives nu.  The amplitude is a
device reduction (``_measure.fields``); only one number per sample leaves the GPU.
"""

from __future__ import annotations

from typing import Any

import numpy as np
import torch

from ..problem import Problem
from . import make_solver
from ._measure import exp_rate, fields, sample, solver_name


def shear_wave_problem(n: int, tau: float, amplitude: float = 0.01, collision: str = "bgk",
                       **kw: Any) -> Problem:
    """Periodic n x n box, rho = 1, u = (A sin(2 pi y / n), 0)."""
    u0 = np.zeros((2, n, n))
    u0[0] += amplitude * np.sin(2 * np.pi * np.arange(n) / n)[:, None]
    return Problem((n, n), tau=tau, collision=collision, periodic=(True, True), init_u=u0,
                   **kw)


def run_shear_wave(tau: float, n: int = 64, amplitude: float = 0.01, steps: int | None = None,
                   samples: int = 20, collision: str = "bgk", backend: str | None = None,
                   variant: str | None = None, factory: Any = None) -> dict:
    p = shear_wave_problem(n, tau, amplitude, collision)
    k2 = (2 * np.pi / n) ** 2
    horizon = steps if steps is not None else int(2.0 / (p.nu * k2))
    s = make_solver(p, backend, variant, factory)
    weight: dict[str, torch.Tensor] = {}

    def mode_amplitude() -> float:
        _, u = fields(s)
        if "w" not in weight:  # 2/n sin(2 pi y / n): projection on the mode
            y = torch.arange(n, dtype=torch.float64, device=u.device)
            weight["w"] = (2.0 / n) * torch.sin(2 * np.pi * y / n)
        return float(u[0].mean(dim=1) @ weight["w"])

    t, amp = sample(s, horizon // samples, samples, mode_amplitude)
    nu = -exp_rate(t, amp) / k2
    return {"tau": tau, "n": n, "nu_theory": p.nu, "nu_measured": nu,
            "rel_error": abs(nu - p.nu) / p.nu, "backend": solver_name(s),
            "t": t.tolist(), "amplitude": amp.tolist()}
