"""The solver interface of the reference (pylbm/backends/base.py:18-103).

A solver advances one :class:`~pylbm_amd.problem.Problem`.  Whatever it stores internally
(precision, AA slot order, device), it exchanges the *canonical state*: the post-collision
populations as a float64 array of shape ``(q, ny, nx)`` or ``(q, nz, ny, nx)``.
"""

from __future__ import annotations

from abc import ABC, abstractmethod
from typing import ClassVar

import numpy as np

from ..problem import Problem


class Unsupported(Exception):
    """A backend/variant cannot run the given problem (or cannot run here)."""


class Solver(ABC):
    # identity and traffic model, as class attributes (pylbm/backends/base.py:27-35)
    backend: ClassVar[str] = ""
    variant: ClassVar[str] = ""
    description: ClassVar[str] = ""
    itemsize: ClassVar[int] = 4        # bytes per stored population value
    step_multiple: ClassVar[int] = 1   # AA solvers advance in pairs of steps
    equivalent: ClassVar[bool] = True  # the 2 q itemsize traffic model applies

    def __init__(self, problem: Problem) -> None:
        ok, why = self.supports(problem)
        if not ok:
            raise Unsupported(f"{self.name}: {why}")
        self.problem = problem
        self.steps_done = 0

    @classmethod
    def available(cls) -> tuple[bool, str]:
        """(True, "") if the solver can run on this machine, else (False, reason)."""
        return True, ""

    @classmethod
    def supports(cls, problem: Problem) -> tuple[bool, str]:
        """(True, "") if the solver implements ``problem``, else (False, reason)."""
        return True, ""

    @abstractmethod
    def set_populations(self, f: np.ndarray) -> None:
        """Replace the state by a canonical one."""

    @abstractmethod
    def populations(self) -> np.ndarray:
        """The canonical state as a float64 numpy array."""

    @abstractmethod
    def macroscopic(self) -> tuple[np.ndarray, np.ndarray]:
        """rho (array shape) and u (d, *array shape), float64, of the post-collision state."""

    @abstractmethod
    def _advance(self, n: int) -> None:
        """Run ``n`` steps; ``n`` is a multiple of ``step_multiple``."""

    def step(self, n: int = 1) -> None:
        if n % self.step_multiple:
            raise ValueError(f"{self.name} advances in multiples of {self.step_multiple} steps")
        self._advance(n)
        self.steps_done += n

    def synchronize(self) -> None:
        """Wait for queued work (nothing to wait for on the host)."""

    @property
    def name(self) -> str:
        return f"{self.backend}/{self.variant}"

    @property
    def bytes_per_node(self) -> int:
        """Minimal traffic of one node update: q populations read, q written."""
        return 2 * self.problem.lattice.q * self.itemsize

    def memory_bytes(self) -> int:
        return self.bytes_per_node * self.problem.n_nodes


def check_box(problem: Problem) -> tuple[bool, str]:
    """Box problems only: no solid nodes, inlet or outlet (pylbm/backends/base.py:99-103)."""
    if problem.is_box():
        return True, ""
    return False, "this variant runs box geometries only (no obstacles, inlet or outlet)"
