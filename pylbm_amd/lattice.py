"""Velocity sets (D2Q9, D3Q19), the D2Q9 MRT basis, relaxation-rate helpers, equilibrium.

Same conventions as the reference (pylbm/lattice.py:1-11): SoA ``f[i, y, x]`` /
``f[i, z, y, x]`` with x fastest; y up; the moving lid is the top wall of the last axis and
moves in +x; walls sit half-way between the last fluid node and the virtual solid node.
The velocity orderings fix the layout of the stored planes and must equal the reference's
(pylbm/lattice.py:49-67); the HIP kernels carry the same tables (csrc/lbm_common.hpp).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

CS2 = 1.0 / 3.0


@dataclass(frozen=True)
class Lattice:
    """A DdQq velocity set (pylbm/lattice.py:22-38)."""

    name: str
    d: int
    c: np.ndarray
    w: np.ndarray
    opp: np.ndarray
    ct: tuple = ()
    wt: tuple = ()
    oppt: tuple = ()

    @property
    def q(self) -> int:
        return len(self.w)


def _lattice(name: str, c: list, w: list) -> Lattice:
    ca = np.asarray(c, dtype=np.int64)
    opp = np.array([next(j for j in range(len(ca)) if (ca[j] == -ci).all()) for ci in ca])
    wa = np.asarray(w, dtype=np.float64)
    return Lattice(name, ca.shape[1], ca, wa, opp,
                   tuple(tuple(int(v) for v in r) for r in ca),
                   tuple(float(v) for v in wa), tuple(int(v) for v in opp))


#: 0 rest, 1 E, 2 N, 3 W, 4 S, 5 NE, 6 NW, 7 SW, 8 SE
D2Q9 = _lattice(
    "D2Q9",
    [[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [-1, 1], [-1, -1], [1, -1]],
    [4 / 9] + [1 / 9] * 4 + [1 / 36] * 4,
)

#: opposite pairs adjacent: (1,2), (3,4), ..., (17,18)
D3Q19 = _lattice(
    "D3Q19",
    [[0, 0, 0],
     [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
     [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0],
     [1, 0, 1], [-1, 0, -1], [1, 0, -1], [-1, 0, 1],
     [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1]],
    [1 / 3] + [1 / 18] * 6 + [1 / 36] * 12,
)

LATTICES = {"D2Q9": D2Q9, "D3Q19": D3Q19}

#: Lallemand & Luo (2000) basis, rows (rho, e, eps, jx, qx, jy, qy, pxx, pxy)
MRT_M_D2Q9 = np.array(
    [[1, 1, 1, 1, 1, 1, 1, 1, 1],
     [-4, -1, -1, -1, -1, 2, 2, 2, 2],
     [4, -2, -2, -2, -2, 1, 1, 1, 1],
     [0, 1, 0, -1, 0, 1, -1, -1, 1],
     [0, -2, 0, 2, 0, 1, -1, -1, 1],
     [0, 0, 1, 0, -1, 1, 1, -1, -1],
     [0, 0, -2, 0, 2, 1, 1, -1, -1],
     [0, 1, -1, 1, -1, 0, 0, 0, 0],
     [0, 0, 0, 0, 0, 1, -1, 1, -1]], dtype=np.float64)
MRT_NORM_D2Q9 = (MRT_M_D2Q9 ** 2).sum(axis=1)


def magic_rate(s_nu: float, magic: float = 3.0 / 16.0) -> float:
    """Rate s_q with Lambda = (1/s_nu - 1/2)(1/s_q - 1/2) = magic (pylbm/lattice.py:90-97)."""
    return 1.0 / (0.5 + magic / (1.0 / s_nu - 0.5))


def mrt_rates(tau: float, s_e: float = 1.64, s_eps: float = 1.54,
              s_q: float | None = None) -> np.ndarray:
    """Diagonal MRT rates in basis order (pylbm/lattice.py:100-109).

    ``s_q`` defaults to the magic rate of s_nu = 1/tau (Lambda = 3/16), as in the reference;
    giving it overrides the energy-flux rate.
    """
    s_nu = 1.0 / tau
    sq = magic_rate(s_nu) if s_q is None else float(s_q)
    return np.array([0.0, s_e, s_eps, 0.0, sq, 0.0, sq, s_nu, s_nu])


def equilibrium(lat: Lattice, rho: np.ndarray, u: np.ndarray) -> np.ndarray:
    """Second-order equilibrium; ``u`` has shape (d, ...) (pylbm/lattice.py:112-119)."""
    usq = sum(u[a] ** 2 for a in range(lat.d))
    out = np.empty((lat.q,) + np.shape(rho), dtype=np.result_type(rho, u))
    for i in range(lat.q):
        cu = sum(lat.c[i, a] * u[a] for a in range(lat.d))
        out[i] = lat.w[i] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq)
    return out
