"""CPU oracle for the fused stream-collide time step (float64 NumPy).

TEST INFRASTRUCTURE ONLY.  Nothing in the product package ``pylbm_amd`` imports this module;
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use
it, and only as the checker (or the CPU baseline that is timed), never as a code path of
the HIP backend.

It restates, in its own code, the unfused algorithm of the reference:

* lattice tables, weights, opposites, MRT basis           pylbm/lattice.py:49-119
* relaxation rates (BGK 1/tau, TRT magic, MRT diagonal)   pylbm/lattice.py:90-109,
                                                          pylbm/problem.py:94-106
* pull streaming with periodic wrap and half-way
  bounce-back, later axes overriding earlier ones,
  Ladd's moving-lid term on the top wall                  pylbm/core.py:88-127
* general geometry: stationary solid nodes (bounce-back
  when the pull source is solid), the velocity inlet on
  the left wall (Ladd term with the inlet profile) and
  the rho = 1 pressure outlet on the right column         pylbm/core.py:56-78,96-148
* moments with the Guo half-force shift                   pylbm/core.py:151-169
* 2nd-order equilibrium                                   pylbm/core.py:172-186
* Guo source term                                         pylbm/core.py:189-200
* BGK / TRT / MRT collision                               pylbm/core.py:203-247
* post-collision macroscopic velocity                     pylbm/core.py:255-266
* canonical state <-> stored layout (16-bit f - w shift,
  AA swapped slots)                                       pylbm/backends/_fused.py:83-103,
                                                          pylbm/codegen.py:461-466
* AA odd/even in-place semantics                          pylbm/codegen.py:10-16,105-328,
                                                          pylbm/backends/cuda/lbm.cu:71-98,165-198

Two extensions the reference does not have are restated here as the checker for the
adjoint work: a per-cell Guo force field ``force_field`` of shape (d, *array_shape) (the
reference only has a uniform force tuple, problem.py:32) and an optional MRT energy-flux
rate override ``s_q`` (the reference hard-wires it to the magic rate, problem.py:102-106;
tests/test_lattice.py:45-46 overrides it by monkeypatching).

How each part is pinned (see tests/test_oracle.py):

* tables: invariants (isotropy, opposites, MRT orthogonality) + the reference's golden
  vectors generated in this container (tests/golden/);
* stream: brute-force per-node rule evaluation on tiny grids; permutation property;
* collide: conservation, MRT(equal rates) == BGK, TRT(s-=s+) == BGK, rest state, golden
  collide vectors from the reference;
* full trajectories: golden vectors of the reference's NumPy solver for the seven
  equivalence cases and the 128^2 cavity after 1000 steps, plus analytic viscosity checks;
* AA semantics: AA pairs == AB steps (exactly, in float64);
* general geometry: golden stream / trajectory vectors of the reference's NumPy solver
  (Schaefer-Turek channel, random solids with inlet, outlet and lid, a 3D obstacle) and a
  brute-force per-node rule;
* per-cell force field: equals the uniform-force path when the field is constant.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# ----------------------------------------------------------------------------- lattices
# Velocity orderings follow pylbm/lattice.py:49-67 (the layout of the stored planes depends
# on them, so they must be identical).
_C2 = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [-1, 1], [-1, -1], [1, -1]])
_W2 = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
_C3 = np.array([
    [0, 0, 0],
    [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
    [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0],
    [1, 0, 1], [-1, 0, -1], [1, 0, -1], [-1, 0, 1],
    [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1],
])
_W3 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)

#: Lallemand-Luo D2Q9 moment basis, rows (rho, e, eps, jx, qx, jy, qy, pxx, pxy)
#: (pylbm/lattice.py:73-86).
MRT_M = np.array([
    [1, 1, 1, 1, 1, 1, 1, 1, 1],
    [-4, -1, -1, -1, -1, 2, 2, 2, 2],
    [4, -2, -2, -2, -2, 1, 1, 1, 1],
    [0, 1, 0, -1, 0, 1, -1, -1, 1],
    [0, -2, 0, 2, 0, 1, -1, -1, 1],
    [0, 0, 1, 0, -1, 1, 1, -1, -1],
    [0, 0, -2, 0, 2, 1, 1, -1, -1],
    [0, 1, -1, 1, -1, 0, 0, 0, 0],
    [0, 0, 0, 0, 0, 1, -1, 1, -1],
], dtype=np.float64)
MRT_NORM = (MRT_M ** 2).sum(axis=1)


@dataclass(frozen=True)
class Lat:
    d: int
    c: np.ndarray
    w: np.ndarray
    opp: np.ndarray

    @property
    def q(self) -> int:
        return len(self.w)


def _mk(c: np.ndarray, w: np.ndarray) -> Lat:
    opp = np.array([int(np.flatnonzero((c == -ci).all(axis=1))[0]) for ci in c])
    return Lat(c.shape[1], c.astype(np.int64), w.astype(np.float64), opp)


D2Q9 = _mk(_C2, _W2)
D3Q19 = _mk(_C3, _W3)


def lattice_for(dim: int) -> Lat:
    return D2Q9 if dim == 2 else D3Q19


def magic_rate(s_nu: float, magic: float = 3.0 / 16.0) -> float:
    """Odd rate from Lambda = (1/s_nu - 1/2)(1/s_q - 1/2)  (pylbm/lattice.py:90-97)."""
    return 1.0 / (0.5 + magic / (1.0 / s_nu - 0.5))


# ----------------------------------------------------------------------------- case
@dataclass
class Case:
    """The subset of a problem the oracle needs (pylbm/problem.py:15-42).

    ``force`` is the uniform Guo force, ``force_field`` (d, *array_shape) a per-cell one
    (mutually exclusive).  ``s_q`` overrides the MRT energy-flux rate.  ``solid`` (bool,
    array layout), ``inlet`` (u_x(y), shape (ny,), 2D) and ``outlet`` (2D) are the general
    geometry of pylbm/problem.py:37-40.
    """

    shape: tuple
    tau: float = 0.6
    collision: str = "bgk"
    periodic: tuple | None = None
    u_lid: float = 0.0
    force: tuple | None = None
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
        d = len(self.shape)
        self.periodic = tuple(bool(p) for p in (self.periodic or (False,) * d))
        self.force = tuple(float(v) for v in (self.force or (0.0,) * d))

    @property
    def dim(self) -> int:
        return len(self.shape)

    @property
    def lat(self) -> Lat:
        return lattice_for(self.dim)

    @property
    def array_shape(self) -> tuple:
        return tuple(reversed(self.shape))

    @property
    def has_force(self) -> bool:
        return self.force_field is not None or any(v != 0.0 for v in self.force)

    def F(self) -> list:
        """Per-axis force (scalars or arrays in array layout)."""
        if self.force_field is not None:
            ff = np.asarray(self.force_field, dtype=np.float64)
            return [ff[a] for a in range(self.dim)]
        return list(self.force)

    def rates(self) -> tuple:
        """BGK: (omega,); TRT: (s+, s-); MRT: the 9 diagonal rates in basis order."""
        s_nu = 1.0 / self.tau
        if self.collision == "bgk":
            return (s_nu,)
        if self.collision == "trt":
            return (s_nu, magic_rate(s_nu, self.magic))
        s_q = magic_rate(s_nu, 3.0 / 16.0) if self.s_q is None else float(self.s_q)
        return (0.0, self.s_e, self.s_eps, 0.0, s_q, 0.0, s_q, s_nu, s_nu)


def equilibrium(lat: Lat, rho: np.ndarray, u) -> np.ndarray:
    """f_i^eq = w_i rho (1 + 3 cu + 4.5 cu^2 - 1.5 u^2)  (pylbm/lattice.py:112-119)."""
    usq = sum(u[a] ** 2 for a in range(lat.d))
    out = np.empty((lat.q,) + np.shape(rho), dtype=np.float64)
    for i in range(lat.q):
        cu = sum(lat.c[i, a] * u[a] for a in range(lat.d))
        out[i] = lat.w[i] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq)
    return out


def initial_populations(case: Case) -> np.ndarray:
    shp = case.array_shape
    rho = np.ones(shp) if case.init_rho is None else np.asarray(case.init_rho, float)
    u = np.zeros((case.dim,) + shp) if case.init_u is None else np.asarray(case.init_u, float)
    return equilibrium(case.lat, rho, u)


# ----------------------------------------------------------------------------- stream
def _face(d: int, axis: int, side: int) -> tuple:
    idx = [slice(None)] * d
    idx[d - 1 - axis] = side
    return tuple(idx)


def stream(f: np.ndarray, case: Case) -> np.ndarray:
    """Pull streaming + half-way bounce-back + Ladd lid (pylbm/core.py:88-127).

    Axis ``a`` of the problem is array axis ``d-1-a``.  Solid sources are bounced back
    first; walls are then applied axis by axis so that a later axis overrides an earlier
    one at edges and corners (the lid, top wall of the last axis, wins at the lid corners,
    and a y wall removes the inlet term at the inlet corners); the outlet overrides
    everything on the last column.
    """
    lat, d = case.lat, case.dim
    axes = tuple(range(d))
    out = np.empty_like(f)
    for i in range(lat.q):
        sh = tuple(int(lat.c[i, a]) for a in reversed(range(d)))
        out[i] = np.roll(f[i], sh, axis=axes) if any(sh) else f[i]
    if case.solid is not None:
        _solid_pull(out, f, case)
    inlet = None
    if case.inlet is not None:
        if d != 2:
            raise ValueError("the inlet is 2D only")
        inlet = np.asarray(case.inlet, dtype=np.float64)
    for a in range(d):
        if case.periodic[a]:
            continue
        for i in range(lat.q):
            ca = lat.c[i, a]
            if ca == 1:  # source x - c_i lies behind the low wall
                fc = _face(d, a, 0)
                val = f[lat.opp[i]][fc]
                if a == 0 and inlet is not None:  # velocity inlet: Ladd term, u_x(y)
                    val = val + 6.0 * lat.w[i] * lat.c[i, 0] * inlet
                out[i][fc] = val
            elif ca == -1:  # source behind the high wall
                fc = _face(d, a, -1)
                val = f[lat.opp[i]][fc]
                if a == d - 1 and case.u_lid != 0.0 and lat.c[i, 0] != 0:
                    val = val + 6.0 * lat.w[i] * lat.c[i, 0] * case.u_lid
                out[i][fc] = val
    if case.outlet:
        _outlet(out, f, case)
    return out


def _solid_pull(out: np.ndarray, f: np.ndarray, case: Case) -> None:
    """Bounce-back where the pull source x - c_i is a solid node (pylbm/core.py:100-104).

    The source is located with a periodic wrap on every axis (np.roll), also on walled
    axes, where the wall rule applied afterwards decides the link anyway.  Solid nodes
    themselves are streamed and collided like any other node.
    """
    lat, d = case.lat, case.dim
    solid = np.asarray(case.solid, dtype=bool)
    for i in range(1, lat.q):
        sh = tuple(int(lat.c[i, a]) for a in reversed(range(d)))
        src_solid = np.roll(solid, sh, axis=tuple(range(d)))
        out[i] = np.where(src_solid, f[lat.opp[i]], out[i])


def _outlet(out: np.ndarray, f: np.ndarray, case: Case) -> None:
    """Pressure outlet (pylbm/core.py:130-148): on the last column the populations with
    c_x = -1 become the rho = 1 equilibrium at the velocity of the stored state one column
    to the left (no force correction)."""
    lat = case.lat
    if case.dim != 2:
        raise ValueError("the outlet is 2D only")
    col = f[:, :, -2]
    rho = col.sum(axis=0)
    ux = (lat.c[:, 0][:, None] * col).sum(axis=0) / rho
    uy = (lat.c[:, 1][:, None] * col).sum(axis=0) / rho
    usq = ux * ux + uy * uy
    for i in range(lat.q):
        if lat.c[i, 0] == -1:
            cu = lat.c[i, 0] * ux + lat.c[i, 1] * uy
            out[i][:, -1] = lat.w[i] * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq)


# ----------------------------------------------------------------------------- collide
def moments(f, case: Case):
    """rho and u = (sum c f + F/2)/rho  (pylbm/core.py:151-169)."""
    lat = case.lat
    F = case.F()
    rho = f[0]
    for i in range(1, lat.q):
        rho = rho + f[i]
    u = []
    for a in range(case.dim):
        j = None
        for i in range(lat.q):
            ci = lat.c[i, a]
            if ci == 0:
                continue
            t = f[i] if ci == 1 else -f[i]
            j = t if j is None else j + t
        if case.force_field is not None or F[a] != 0.0:
            j = j + 0.5 * F[a]
        u.append(j / rho)
    return rho, u


def collide(f, case: Case) -> np.ndarray:
    """BGK / TRT / MRT with Guo forcing (pylbm/core.py:172-247)."""
    lat, d = case.lat, case.dim
    rho, u = moments(f, case)
    F = case.F()
    usq = sum(u[a] * u[a] for a in range(d))
    cus, feq = [], []
    for i in range(lat.q):
        cu = sum(lat.c[i, a] * u[a] for a in range(d))
        cus.append(cu)
        feq.append(lat.w[i] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq))
    src = None
    if case.has_force:
        uf = sum(u[a] * F[a] for a in range(d))
        src = []
        for i in range(lat.q):
            cf = sum(lat.c[i, a] * F[a] for a in range(d))
            src.append(lat.w[i] * (3.0 * (cf - uf) + 9.0 * cus[i] * cf))
    r = case.rates()
    out = []
    if case.collision == "bgk":
        om = r[0]
        for i in range(lat.q):
            v = f[i] - om * (f[i] - feq[i])
            if src is not None:
                v = v + (1.0 - 0.5 * om) * src[i]
            out.append(v)
    elif case.collision == "trt":
        sp, sm = r
        for i in range(lat.q):
            o = lat.opp[i]
            v = f[i] - sp * (0.5 * (f[i] + f[o]) - 0.5 * (feq[i] + feq[o])) \
                - sm * (0.5 * (f[i] - f[o]) - 0.5 * (feq[i] - feq[o]))
            if src is not None:
                v = v + (1.0 - 0.5 * sp) * 0.5 * (src[i] + src[o]) \
                    + (1.0 - 0.5 * sm) * 0.5 * (src[i] - src[o])
            out.append(v)
    elif case.collision == "mrt":
        if d != 2:
            raise ValueError("MRT is D2Q9 only")
        dm = []
        for k in range(lat.q):
            acc = 0.0
            for i in range(lat.q):
                m = MRT_M[k, i]
                if m == 0.0:
                    continue
                t = -r[k] * m * (f[i] - feq[i])
                if src is not None:
                    t = t + (1.0 - 0.5 * r[k]) * m * src[i]
                acc = acc + t
            dm.append(acc)
        for i in range(lat.q):
            v = f[i]
            for k in range(lat.q):
                if MRT_M[k, i] != 0.0:
                    v = v + (MRT_M[k, i] / MRT_NORM[k]) * dm[k]
            out.append(v)
    else:
        raise ValueError(case.collision)
    return np.stack([np.broadcast_to(np.asarray(v, dtype=np.float64), np.shape(f[0]))
                     for v in out])


def step(f: np.ndarray, case: Case) -> np.ndarray:
    return collide(stream(f, case), case)


def run(case: Case, steps: int, f0: np.ndarray | None = None) -> np.ndarray:
    f = initial_populations(case) if f0 is None else np.array(f0, dtype=np.float64)
    for _ in range(steps):
        f = step(f, case)
    return f


def macroscopic(f, case: Case):
    """(rho, u) of a post-collision state: u = (sum c f* - F/2)/rho (pylbm/core.py:255-266)."""
    rho, u = moments([f[i] for i in range(case.lat.q)], case)
    if case.has_force:
        F = case.F()
        u = [(u[a] * rho - F[a]) / rho for a in range(case.dim)]
    return rho, np.stack([np.broadcast_to(c, np.shape(rho)) for c in u])


class OracleSolver:
    """Minimal stepping object with the Solver-API shape, for the analytic checks on CPU."""

    step_multiple = 1
    name = "oracle/numpy-f64"

    def __init__(self, case: Case) -> None:
        self.case = case
        self.f = initial_populations(case)
        self.steps_done = 0

    def step(self, n: int = 1) -> None:
        for _ in range(n):
            self.f = step(self.f, self.case)
        self.steps_done += n

    def populations(self) -> np.ndarray:
        return self.f.copy()

    def macroscopic(self):
        return macroscopic(self.f, self.case)

    def synchronize(self) -> None:
        return None


# ----------------------------------------------------------------------------- storage
def bf16_rne_bits(x32: np.ndarray) -> np.ndarray:
    """float32 -> bfloat16 bit pattern, round to nearest, ties to even (finite inputs)."""
    u = np.ascontiguousarray(x32, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = u + 0x7FFF + ((u >> 16) & 1)
    return (u >> 16).astype(np.uint16)


def bf16_bits_to_f64(b: np.ndarray) -> np.ndarray:
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32).astype(np.float64)


def encode(f: np.ndarray, storage: str, aa: bool, lat: Lat) -> np.ndarray:
    """Canonical float64 state -> stored array (pylbm/backends/_fused.py:83-90).

    16-bit formats store f - w_i, rounded float64 -> float32 -> 16 bit; the AA layout keeps
    population i in slot opp(i).  bf16 is returned as its uint16 bit pattern.
    """
    g = np.asarray(f, dtype=np.float64)
    if storage != "f32":
        g = g - lat.w.reshape((-1,) + (1,) * (g.ndim - 1))
    if aa:
        g = g[lat.opp]
    g32 = g.astype(np.float32)
    if storage in ("f32", "f32s"):  # f32s: shifted fp32 (no reference counterpart)
        return g32
    if storage == "f16":
        return g32.astype(np.float16)
    return bf16_rne_bits(g32)


def decode(s: np.ndarray, storage: str, aa: bool, lat: Lat) -> np.ndarray:
    """Stored array -> canonical float64 state (pylbm/backends/_fused.py:95-103)."""
    g = bf16_bits_to_f64(s) if storage == "bf16" else np.asarray(s).astype(np.float64)
    if aa:
        g = g[lat.opp]
    if storage != "f32":
        g = g + lat.w.reshape((-1,) + (1,) * (g.ndim - 1))
    return g


# ----------------------------------------------------------------------------- AA pattern
def _neigh(case: Case, i: int, sign: int):
    """For every node: flat index of x + sign*c_i (periodic wrap), wall mask, lid mask."""
    lat, d = case.lat, case.dim
    ashape = case.array_shape
    grids = np.meshgrid(*[np.arange(s) for s in ashape], indexing="ij")
    coord = [grids[d - 1 - a] for a in range(d)]  # problem axis a -> coordinate array
    tgt = []
    wall = np.zeros(ashape, dtype=bool)
    top = np.zeros(ashape, dtype=bool)
    for a in range(d):
        t = coord[a] + sign * int(lat.c[i, a])
        out = (t < 0) | (t >= case.shape[a])
        if case.periodic[a]:
            t = t % case.shape[a]
        else:
            wall |= out
            if a == d - 1:
                top |= t >= case.shape[a]
            t = np.clip(t, 0, case.shape[a] - 1)
        tgt.append(t)
    flat = np.zeros(ashape, dtype=np.int64)
    stride = 1
    for a in range(d):
        flat += tgt[a] * stride
        stride *= case.shape[a]
    return flat, wall, top


def aa_odd(A: np.ndarray, case: Case) -> np.ndarray:
    """One AA odd step on a float64 array in the swapped layout (pylbm/codegen.py:10-16).

    Reads A[opp(i), x - c_i] (own A[i, x] on a wall link, plus the lid term), collides,
    writes A[i, x + c_i] (own A[opp(i), x] on a wall link, minus the lid term).
    """
    lat = case.lat
    q, N = lat.q, int(np.prod(case.shape))
    flatA = A.reshape(q, N)
    own = np.arange(N)
    f = np.empty((q, N))
    for i in range(q):
        src, wall, top = _neigh(case, i, -1)
        src, wall, top = src.reshape(-1), wall.reshape(-1), top.reshape(-1)
        v = np.where(wall, flatA[i, own], flatA[lat.opp[i], src])
        if case.u_lid != 0.0 and lat.c[i, 0] != 0:
            v = v + np.where(wall & top, 6.0 * lat.w[i] * lat.c[i, 0] * case.u_lid, 0.0)
        f[i] = v
    p = collide(f.reshape((q,) + case.array_shape), case).reshape(q, N)
    outA = flatA.copy()
    for i in range(q):
        dst, wall, top = _neigh(case, i, +1)
        dst, wall, top = dst.reshape(-1), wall.reshape(-1), top.reshape(-1)
        corr = 0.0
        if case.u_lid != 0.0 and lat.c[i, 0] != 0:
            corr = np.where(wall & top, -6.0 * lat.w[i] * lat.c[i, 0] * case.u_lid, 0.0)
        bb = p[i] + corr
        outA[lat.opp[i], own[wall]] = bb[wall] if np.ndim(bb) else bb
        outA[i, dst[~wall]] = p[i][~wall]
    return outA.reshape(A.shape)


def aa_even(A: np.ndarray, case: Case) -> np.ndarray:
    """One AA even step: read A[i, x], collide, write A[opp(i), x]."""
    lat = case.lat
    p = collide(np.asarray(A, dtype=np.float64), case)
    return p[lat.opp]
