"""Velocity sets (D2Q9, D3Q19), the D2Q9 MRT basis, relaxation-rate helpers, equilibrium.

Conventions of the reference (pylbm/lattice.py:1-11): SoA ``f[i, y, x]`` / ``f[i, z, y, x]``
with x fastest; y up; the moving lid is the top wall of the last axis and moves in +x;
walls sit half-way between the last fluid node and the virtual solid node.  The velocity
orderings fix the layout of the stored planes and equal the reference's
(pylbm/lattice.py:49-67); the HIP kernels carry the same tables (csrc/lbm_common.hpp).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass(frozen=True)
class Lattice:
    """A DdQq velocity set: integer velocities ``c`` (q, d) and weights ``w`` (q,); the
    dimension, population count and opposite-direction table follow from them."""

    name: str
    c: np.ndarray
    w: np.ndarray
    opp: np.ndarray = field(init=False)

    def __post_init__(self) -> None:
        c = np.asarray(self.c, dtype=np.int64)
        w = np.asarray(self.w, dtype=np.float64)
        # opposite of i: the row of -c_i (a row-wise match against the whole table)
        match = (c[None, :, :] == -c[:, None, :]).all(axis=2)
        object.__setattr__(self, "c", c)
        object.__setattr__(self, "w", w)
        object.__setattr__(self, "opp", match.argmax(axis=1))

    @property
    def d(self) -> int:
        return self.c.shape[1]

    @property
    def q(self) -> int:
        return self.c.shape[0]


#: 0 rest, 1 E, 2 N, 3 W, 4 S, 5 NE, 6 NW, 7 SW, 8 SE
D2Q9 = Lattice("D2Q9",
               [[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [-1, 1], [-1, -1], [1, -1]],
               [4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)

#: opposite pairs adjacent: (1,2), (3,4), ..., (17,18)
D3Q19 = Lattice("D3Q19",
                [[0, 0, 0],
                 [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                 [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0],
                 [1, 0, 1], [-1, 0, -1], [1, 0, -1], [-1, 0, 1],
                 [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1]],
                [1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)

LATTICES = {lat.name: lat for lat in (D2Q9, D3Q19)}

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
#: squared row norms: the basis is orthogonal, so M^-1 = M^T diag(1 / norm)
MRT_NORM_D2Q9 = np.einsum("ki,ki->k", MRT_M_D2Q9, MRT_M_D2Q9)


def magic_rate(s_nu: float, magic: float = 3.0 / 16.0) -> float:
    """The odd-moment rate s_q for which (1/s_nu - 1/2)(1/s_q - 1/2) = magic
    (pylbm/lattice.py:90-97; Lambda = 3/16 puts the half-way wall exactly mid-link)."""
    tau_odd = 0.5 + magic / (1.0 / s_nu - 0.5)
    return 1.0 / tau_odd


def mrt_rates(tau: float, s_e: float = 1.64, s_eps: float = 1.54,
              s_q: float | None = None) -> np.ndarray:
    """Diagonal MRT rates in basis order (pylbm/lattice.py:100-109): 0 for the conserved
    moments, s_e, s_eps, the energy-flux rate s_q (the magic rate of 1/tau unless given) and
    the shear rate 1/tau."""
    s_nu = 1.0 / tau
    sq = magic_rate(s_nu) if s_q is None else float(s_q)
    rates = np.zeros(9)
    rates[[1, 2]] = s_e, s_eps
    rates[[4, 6]] = sq
    rates[[7, 8]] = s_nu
    return rates


def equilibrium(lat: Lattice, rho: np.ndarray, u: np.ndarray) -> np.ndarray:
    """Second-order equilibrium w_i rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u^2) for every
    population; ``u`` has shape (d, ...) (pylbm/lattice.py:112-119)."""
    u = np.asarray(u)
    rho = np.asarray(rho)
    usq = (u * u).sum(axis=0)
    cu = np.tensordot(lat.c.astype(np.float64), u, axes=(1, 0))  # (q, ...)
    w = lat.w.reshape((-1,) + (1,) * rho.ndim)
    return w * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq)
