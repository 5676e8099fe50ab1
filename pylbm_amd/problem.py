"""Problem description shared by every solver (mirrors pylbm/problem.py:15-126).

Two additions over the reference, both needed by the differentiable runs:

* ``force_field``: a per-cell Guo body force of shape (d, *array_shape) (the reference only
  has the uniform ``force`` tuple, pylbm/problem.py:32);
* ``s_q``: an optional MRT energy-flux rate (the reference fixes it to the magic rate of
  1/tau, pylbm/problem.py:102-106).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .lattice import D2Q9, D3Q19, Lattice, equilibrium, magic_rate, mrt_rates

COLLISIONS = ("bgk", "trt", "mrt")


@dataclass
class Problem:
    """A lattice-Boltzmann problem on a rectangular box.

    ``shape`` is (nx, ny) or (nx, ny, nz).  Each axis is periodic or bounded by half-way
    bounce-back walls; the top wall of the last axis moves with ``u_lid`` in +x (Ladd
    correction, wall density 1).  ``solid`` (bool, array layout), ``inlet`` (u_x(y) on the
    x = 0 wall, 2D) and ``outlet`` (rho = 1 on the last column, 2D) describe general
    geometry (pylbm/problem.py:37-40); the hip two-buffer variants run it, the AA variants
    take box problems only (:meth:`is_box`).
    """

    shape: tuple[int, ...]
    tau: float = 0.6
    collision: str = "bgk"
    periodic: tuple[bool, ...] | None = None
    u_lid: float = 0.0
    force: tuple[float, ...] | None = None
    magic: float = 3.0 / 16.0
    s_e: float = 1.64
    s_eps: float = 1.54
    s_q: float | None = None
    force_field: np.ndarray | None = None
    solid: np.ndarray | None = None
    inlet: np.ndarray | None = None
    outlet: bool = False
    init_rho: np.ndarray | None = None
    init_u: np.ndarray | None = None
    meta: dict = field(default_factory=dict)

    def __post_init__(self) -> None:
        self.shape = tuple(int(s) for s in self.shape)
        if self.dim not in (2, 3):
            raise ValueError("shape must be 2D or 3D")
        if any(s < 1 for s in self.shape):
            raise ValueError("every extent must be >= 1")
        if self.periodic is None:
            self.periodic = (False,) * self.dim
        self.periodic = tuple(bool(p) for p in self.periodic)
        if len(self.periodic) != self.dim:
            raise ValueError("periodic needs one flag per axis")
        if self.force is None:
            self.force = (0.0,) * self.dim
        self.force = tuple(float(v) for v in self.force)
        if len(self.force) != self.dim:
            raise ValueError("force needs one component per axis")
        if self.force_field is not None:
            ff = np.asarray(self.force_field, dtype=np.float64)
            if ff.shape != (self.dim,) + self.array_shape:
                raise ValueError(f"force_field must have shape {(self.dim,) + self.array_shape}")
            if any(v != 0.0 for v in self.force):
                raise ValueError("give either a uniform force or a force_field, not both")
            self.force_field = ff
        if self.solid is not None:
            sol = np.asarray(self.solid, dtype=bool)
            if sol.shape != self.array_shape:
                raise ValueError(f"solid must have shape {self.array_shape}")
            self.solid = sol
        if self.inlet is not None:
            if self.dim != 2:
                raise ValueError("inlet only implemented in 2D")
            inl = np.asarray(self.inlet, dtype=np.float64)
            if inl.shape != (self.shape[1],):
                raise ValueError(f"inlet must have shape ({self.shape[1]},)")
            self.inlet = inl
        if self.outlet and self.dim != 2:
            raise ValueError("outlet only implemented in 2D")
        if self.collision not in COLLISIONS:
            raise ValueError(f"collision must be one of {COLLISIONS}")
        if self.collision == "mrt" and self.dim != 2:
            raise ValueError("MRT is implemented for D2Q9 only")
        if self.u_lid and self.periodic[-1]:
            raise ValueError("a moving lid needs a wall on the last axis")
        if not 0.5 < self.tau:
            raise ValueError("tau must be > 0.5")

    # ------------------------------------------------------------------ properties
    @property
    def dim(self) -> int:
        return len(self.shape)

    @property
    def lattice(self) -> Lattice:
        return D2Q9 if self.dim == 2 else D3Q19

    @property
    def array_shape(self) -> tuple[int, ...]:
        """(ny, nx) or (nz, ny, nx)."""
        return tuple(reversed(self.shape))

    @property
    def n_nodes(self) -> int:
        return int(np.prod(self.shape))

    @property
    def nu(self) -> float:
        return (self.tau - 0.5) / 3.0

    @property
    def has_force(self) -> bool:
        return self.force_field is not None or any(v != 0.0 for v in self.force)

    def is_box(self) -> bool:
        return self.solid is None and self.inlet is None and not self.outlet

    def trt_rates(self) -> tuple[float, float]:
        s_plus = 1.0 / self.tau
        return s_plus, magic_rate(s_plus, self.magic)

    def mrt_rates(self) -> np.ndarray:
        return mrt_rates(self.tau, self.s_e, self.s_eps, self.s_q)

    def mrt_rate_tuple(self) -> tuple[float, ...]:
        return tuple(float(v) for v in self.mrt_rates())

    def runtime_rates(self) -> tuple[float, float, float, float]:
        """(r0, r1, r2, r3) of the kernels (pylbm/codegen.py:437-445)."""
        if self.collision == "bgk":
            return 1.0 / self.tau, 0.0, 0.0, 0.0
        if self.collision == "trt":
            sp, sm = self.trt_rates()
            return sp, sm, 0.0, 0.0
        s = self.mrt_rates()
        return float(s[7]), float(s[1]), float(s[2]), float(s[4])

    def initial_populations(self) -> np.ndarray:
        """Equilibrium populations of the initial fields (float64, SoA)."""
        shp = self.array_shape
        rho = np.ones(shp) if self.init_rho is None else np.asarray(self.init_rho, float)
        u = np.zeros((self.dim,) + shp) if self.init_u is None else np.asarray(self.init_u, float)
        return equilibrium(self.lattice, rho, u)

    @staticmethod
    def cavity(n: int | tuple[int, ...], re: float | None = None, u_lid: float = 0.1,
               tau: float | None = None, collision: str = "bgk", dim: int = 2) -> Problem:
        """Lid-driven cavity; give ``re`` (Re = u_lid * nx / nu) or ``tau``."""
        shape = (n,) * dim if isinstance(n, int) else tuple(n)
        if tau is None:
            if re is None:
                raise ValueError("give re or tau")
            tau = 3.0 * u_lid * shape[0] / re + 0.5
        return Problem(shape, tau=tau, collision=collision, u_lid=u_lid)
