"""Measurement core of the validation cases: observables, sampling, fits, steady states.

Every observable is computed where the state lives.  A hip solver exposes its float64
density and velocity as device tensors (``macroscopic_device``), so a decay amplitude, a
channel profile or the change between two snapshots is a small device reduction and only
scalars or single lines cross to the host; a 512^3 state is never copied out.  Any other
solver (the float64 oracle on the CPU, for instance) only needs ``macroscopic()`` and is
measured on host tensors with the same code.
"""

from __future__ import annotations

from typing import Any, Callable

import numpy as np
import torch


def fields(solver: Any) -> tuple[torch.Tensor, torch.Tensor]:
    """(rho, u) of the post-collision state as float64 tensors, on the solver's device."""
    dev = getattr(solver, "macroscopic_device", None)
    if dev is not None:
        return dev()
    rho, u = solver.macroscopic()
    return torch.as_tensor(np.asarray(rho, dtype=np.float64)), \
        torch.as_tensor(np.asarray(u, dtype=np.float64))


def mean_energy(solver: Any) -> float:
    """Mean kinetic energy per node, 0.5 rho |u|^2 (a float64 device reduction if offered)."""
    red = getattr(solver, "reduce", None)
    if red is not None:
        return red("kinetic") / solver.problem.n_nodes
    rho, u = fields(solver)
    return float(0.5 * (rho * (u * u).sum(dim=0)).mean())


def sample(solver: Any, interval: int, count: int,
           observe: Callable[[], float]) -> tuple[np.ndarray, np.ndarray]:
    """``observe()`` at steps 0, interval, ..., count * interval from the current state
    (``interval`` rounded down to the solver's step multiple)."""
    m = getattr(solver, "step_multiple", 1)
    interval = max(m, interval - interval % m)
    t, v = np.empty(count + 1), np.empty(count + 1)
    for j in range(count + 1):
        if j:
            solver.step(interval)
        t[j] = solver.steps_done
        v[j] = observe()
    return t, v


def exp_rate(t: np.ndarray, v: np.ndarray) -> float:
    """Least-squares rate r of |v| ~ exp(r t) (closed-form regression of log |v| on t)."""
    t = np.asarray(t, dtype=np.float64)
    y = np.log(np.abs(np.asarray(v, dtype=np.float64)))
    dt = t - t.mean()
    return float((dt * (y - y.mean())).sum() / (dt * dt).sum())


class Steady:
    """Steady-state monitor fed with the change between two checks.

    ``update(change)`` returns True once the change falls below ``tol``, "plateau" once it
    has not improved by the factor ``improve`` for ``patience`` checks while its best value
    is below ``floor`` (16-bit and fp32 round-off stop the decay above ``tol``), else None.
    """

    def __init__(self, tol: float, improve: float, patience: int, floor: float) -> None:
        self.tol, self.improve, self.patience, self.floor = tol, improve, patience, floor
        self.best, self.stale = np.inf, 0

    def update(self, change: float) -> bool | str | None:
        if change < self.tol:
            return True
        if change < self.improve * self.best:
            self.best, self.stale = change, 0
        else:
            self.stale += 1
        if self.stale >= self.patience and self.best < self.floor:
            return "plateau"
        return None


def solver_name(solver: Any) -> str:
    return getattr(solver, "name", type(solver).__name__)
