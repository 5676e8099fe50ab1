"""The ``hip`` backend: hand-written gfx950 fused kernels behind the Solver API.

Mirrors the reference's FusedSolver contract (pylbm/backends/_fused.py:42-136) with a new
native rung instead of CuPy/NVRTC (pylbm/backends/cupy_backend.py:76-122):

* device memory is torch-owned (``torch.empty(q*N, device="cuda")``); the kernels borrow
  raw pointers through the C ABI of libpylbm_hip.so (include/pylbm_hip.h);
* every launch goes to torch's current HIP stream;
* ``_advance`` enqueues the whole step loop from C (``pylbm_hip_run``), one call per
  ``step(n)``, instead of one Python launch per step;
* encode / decode / initial equilibrium / observables run on the device, so a 512^3 D3Q19
  problem never materialises its 20 GB float64 state on the host;
* the two-buffer variants also run general geometry (solid nodes, velocity inlet, pressure
  outlet: pylbm/core.py:56-148), which the reference's fused kernels reject; ``gather`` is
  the fp32 entry the reference's Karman runner asks for;
* there is no fallback: without the library or a GPU the entries report unavailable and
  ``make`` raises.
"""

from __future__ import annotations

import ctypes
from typing import Any, ClassVar

import numpy as np

from .. import _lib
from ..problem import Problem
from .base import Solver, check_box

try:
    import torch
except Exception:  # pragma: no cover
    torch = None  # type: ignore[assignment]

_TORCH_DTYPES = {"f32": "float32", "f32s": "float32", "f16": "float16", "bf16": "bfloat16"}
#: the kernels index a plane with 32-bit node indices (one row segment of slack) and 32-bit
#: byte offsets, so a plane holds at most 4 GiB
MAX_NODES = 2**31 - 1 - 256


def gpu_ok() -> tuple[bool, str]:
    if torch is None:
        return False, "torch not installed"
    err = _lib.load_error()
    if err:
        return False, err
    if not torch.cuda.is_available():
        return False, "no HIP device"
    return True, ""


def _ptr(t: Any) -> int | None:
    return None if t is None else int(t.data_ptr())


def _stream() -> int:
    return int(torch.cuda.current_stream().cuda_stream)


class HipFused(Solver):
    """gfx950 fused stream-collide, two buffers, fp32 populations."""

    backend = "hip"
    variant = "fused"
    description = "hand-written HIP (gfx950), fused pull, two buffers, fp32"
    storage: ClassVar[str] = "f32"
    aa: ClassVar[bool] = False
    #: two time steps per launch (the intermediate state on chip; D2Q9 boxes)
    pair: ClassVar[bool] = False
    lattices: ClassVar[tuple[str, ...]] = ("D2Q9", "D3Q19")
    collisions: ClassVar[tuple[str, ...]] = ("bgk", "trt", "mrt")

    @classmethod
    def available(cls) -> tuple[bool, str]:
        return gpu_ok()

    @classmethod
    def supports(cls, problem: Problem) -> tuple[bool, str]:
        item = 2 if cls.storage in ("f16", "bf16") else 4
        if not problem.is_box():
            # obstacles, inlet and outlet: two-buffer kernels, whole array under 4 GiB
            if cls.aa:
                return check_box(problem)
            if problem.lattice.q * problem.n_nodes * item >= 2**32:
                return False, "with obstacles, inlet or outlet the state must be < 4 GiB"
        if problem.lattice.name not in cls.lattices:
            return False, f"{problem.lattice.name} not implemented"
        if cls.pair and not problem.is_box():
            return False, "two steps per launch run on box problems (no obstacles, inlet or outlet)"
        if cls.pair and problem.lattice.q * problem.n_nodes * item >= 2**32:
            return False, "two steps per launch need the state under 4 GiB"
        if problem.collision not in cls.collisions:
            return False, f"{problem.collision} not implemented"
        if problem.n_nodes > MAX_NODES or problem.n_nodes * item >= 2**32:
            return False, "a population plane exceeds the kernels' 32-bit offsets"
        return True, ""

    def __init__(self, problem: Problem, init: str = "auto") -> None:
        super().__init__(problem)
        self.step_multiple = 2 if (self.aa or self.pair) else 1
        self.itemsize = 2 if self.storage in ("f16", "bf16") else 4
        self.q = problem.lattice.q
        self.N = problem.n_nodes
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.tdtype = getattr(torch, _TORCH_DTYPES[self.storage])
        self.force_field = None
        if problem.force_field is not None:
            self.force_field = torch.as_tensor(
                np.ascontiguousarray(problem.force_field, dtype=np.float32).reshape(
                    problem.dim, -1), device=self.device)
        self.A = torch.empty(self.q * self.N, device=self.device, dtype=self.tdtype)
        self.B = None if self.aa else torch.empty_like(self.A)
        self.links = self.link_groups = self.inlet = None
        if problem.inlet is not None:
            self.inlet = torch.as_tensor(np.asarray(problem.inlet, dtype=np.float32),
                                         device=self.device)
        self.params = self._make_params()
        if problem.solid is not None:
            self.set_solid(problem.solid)
        if init == "auto":
            self.init_equilibrium(problem.init_rho, problem.init_u)
        elif init != "none":
            raise ValueError("init must be 'auto' or 'none'")

    # ------------------------------------------------------------------ parameters
    def _make_params(self) -> _lib.Params:
        p = self.problem
        prm = _lib.Params()
        prm.lattice = _lib.LATTICE_IDS[p.lattice.name]
        prm.collision = _lib.COLLISION_IDS[p.collision]
        prm.storage = _lib.STORAGE_IDS[self.storage]
        prm.aa = 1 if self.aa else 0
        per = tuple(p.periodic) + (True,) * (3 - p.dim)
        shp = tuple(p.shape) + (1,) * (3 - p.dim)
        for k in range(3):
            prm.periodic[k] = 1 if per[k] else 0
        prm.nx, prm.ny, prm.nz = shp
        prm.n_nodes = self.N
        prm.ulid = float(p.u_lid)
        force = tuple(p.force) + (0.0,) * (3 - p.dim)
        for k in range(3):
            prm.force[k] = force[k]
        for k, r in enumerate(p.runtime_rates()):
            prm.rates[k] = r
        prm.force_field = _ptr(self.force_field)
        prm.links = _ptr(self.links)
        prm.link_groups = _ptr(self.link_groups)
        prm.inlet = _ptr(self.inlet)
        prm.outlet = 1 if p.outlet else 0
        return prm

    def set_solid(self, solid: Any) -> None:
        """Install a solid field (bool, array layout) and build its per-node link mask on
        the device (pylbm/core.py:56-71)."""
        st = torch.as_tensor(np.ascontiguousarray(np.asarray(solid, dtype=bool)).astype(np.uint8)
                             if not torch.is_tensor(solid) else solid.to(torch.uint8),
                             device=self.device).contiguous()
        if st.numel() != self.N:
            raise ValueError(f"solid must have shape {self.problem.array_shape}")
        dt = torch.int16 if self.q == 9 else torch.int32
        links = torch.empty(self.N, dtype=dt, device=self.device)
        _lib.call("pylbm_hip_links", self.pref, _ptr(st), _ptr(links), _stream())
        # one flag per 64-node group of a row: the waves of groups that touch a solid (or an
        # x wall) load their bounce-back values with the pull, the others skip the mask
        p = self.problem
        nx = p.shape[0]
        rows = self.N // nx
        groups = torch.empty(rows * ((nx + 63) // 64), dtype=torch.int32, device=self.device)
        _lib.call("pylbm_hip_link_groups", self.pref, _ptr(links), _ptr(groups), _stream())
        torch.cuda.current_stream().synchronize()
        self.links, self.link_groups = links, groups
        self.params.links = _ptr(links)
        self.params.link_groups = _ptr(groups)

    def set_force_field(self, field: Any) -> None:
        """Replace the per-cell force (d, *array_shape) without rebuilding the solver.

        Accepts a numpy array or a float32 device tensor; only valid when the problem was
        created with a ``force_field``.
        """
        if self.force_field is None:
            raise ValueError("problem has no force_field")
        t = torch.as_tensor(field, dtype=torch.float32, device=self.device).reshape(
            self.problem.dim, -1)
        self.force_field.copy_(t)

    @property
    def pref(self) -> Any:
        return ctypes.byref(self.params)

    # ------------------------------------------------------------------ state
    def init_equilibrium(self, rho: Any = None, u: Any = None) -> None:
        """Equilibrium state from rho (array_shape) and u (d, *array_shape), on the device.

        Either may be None (rho = 1, u = 0); numpy arrays or torch tensors are accepted and
        are used in float64, exactly as Problem.initial_populations + set_populations.
        """
        rt = None if rho is None else torch.as_tensor(rho, dtype=torch.float64,
                                                      device=self.device).contiguous()
        ut = None if u is None else torch.as_tensor(u, dtype=torch.float64,
                                                    device=self.device).contiguous()
        if rt is not None and rt.numel() != self.N:
            raise ValueError("rho has the wrong size")
        if ut is not None and ut.numel() != self.N * self.problem.dim:
            raise ValueError("u has the wrong size")
        _lib.call("pylbm_hip_init_equilibrium", self.pref, _ptr(rt), _ptr(ut),
                  _ptr(self.A), _stream())
        torch.cuda.current_stream().synchronize()

    def set_populations(self, f: Any) -> None:
        """Load a canonical float64 state (pylbm/backends/_fused.py:88-89)."""
        t = torch.as_tensor(np.asarray(f, dtype=np.float64) if not torch.is_tensor(f) else f,
                            dtype=torch.float64, device=self.device).contiguous()
        if t.numel() != self.q * self.N:
            raise ValueError(f"expected {(self.q,) + self.problem.array_shape} populations")
        _lib.call("pylbm_hip_encode", self.pref, _ptr(t), _ptr(self.A), _stream())
        torch.cuda.current_stream().synchronize()

    def populations_device(self) -> Any:
        """Canonical float64 state as a (q, *array_shape) device tensor."""
        out = torch.empty((self.q,) + self.problem.array_shape, dtype=torch.float64,
                          device=self.device)
        _lib.call("pylbm_hip_decode", self.pref, _ptr(self.A), _ptr(out), _stream())
        return out

    def populations(self) -> np.ndarray:
        return self.populations_device().cpu().numpy()

    def macroscopic_device(self) -> tuple[Any, Any]:
        shp = self.problem.array_shape
        rho = torch.empty(shp, dtype=torch.float64, device=self.device)
        u = torch.empty((self.problem.dim,) + shp, dtype=torch.float64, device=self.device)
        _lib.call("pylbm_hip_macroscopic", self.pref, _ptr(self.A), _ptr(rho), _ptr(u),
                  _stream())
        return rho, u

    def macroscopic(self) -> tuple[np.ndarray, np.ndarray]:
        rho, u = self.macroscopic_device()
        return rho.cpu().numpy(), u.cpu().numpy()

    def reduce(self, kind: str, other: Any = None) -> float:
        """Float64 reduction on the device: 'mass', 'kinetic', 'maxabs_diff' or 'dot' (with
        ``other`` a solver with the same parameters), or 'loss' (0.5 sum |u - u_target|^2,
        ``other`` the target velocity (d, *array_shape))."""
        k = {"mass": _lib.RED_MASS, "kinetic": _lib.RED_KINETIC,
             "maxabs_diff": _lib.RED_MAXABS_DIFF, "loss": _lib.RED_LOSS,
             "dot": _lib.RED_DOT}[kind]
        second = None
        if kind == "loss":
            second = torch.as_tensor(other, dtype=torch.float64, device=self.device).contiguous()
            if second.numel() != self.problem.dim * self.N:
                raise ValueError("the target velocity has the wrong size")
            ptr = _ptr(second)
        else:
            ptr = _ptr(other.A) if other is not None else None
        out = torch.empty(1, dtype=torch.float64, device=self.device)
        _lib.call("pylbm_hip_reduce", self.pref, k, _ptr(self.A), ptr, _ptr(out), _stream())
        return float(out.item())

    def memory_bytes(self) -> int:
        return (1 if self.aa else 2) * self.q * self.N * self.itemsize

    # ------------------------------------------------------------------ stepping
    def launch(self, pattern: str) -> None:
        """One kernel launch: 'ab' (A -> B, then swap), 'ab2' (two steps A -> B, then swap;
        D2Q9 boxes), 'aa-odd' or 'aa-even' (in place)."""
        pid = _lib.PATTERN_IDS[pattern]
        if pid in (_lib.AB, _lib.AB2):
            _lib.call("pylbm_hip_step", self.pref, _ptr(self.A), _ptr(self.B), pid, _stream())
            self.A, self.B = self.B, self.A
        else:
            _lib.call("pylbm_hip_step", self.pref, _ptr(self.A), _ptr(self.A), pid, _stream())

    def _advance(self, n: int) -> None:
        if self.pair:
            _lib.call("pylbm_hip_run_pairs", self.pref, _ptr(self.A), _ptr(self.B), int(n),
                      _stream())
            if (n // 2) % 2:
                self.A, self.B = self.B, self.A
            return
        _lib.call("pylbm_hip_run", self.pref, _ptr(self.A), _ptr(self.B), int(n), _stream())
        if not self.aa and n % 2:
            self.A, self.B = self.B, self.A

    def synchronize(self) -> None:
        torch.cuda.synchronize()


class HipGather(HipFused):
    """Any geometry the reference's gather solvers run (obstacles, velocity inlet,
    pressure outlet; pylbm/backends/torch_backend.py:106-137) as one fused gfx950 kernel:
    the two-buffer kernel plus a per-node solid-source bit mask (no q*N index table)."""

    variant = "gather"
    description = "hand-written HIP (gfx950), fused pull with obstacles / inlet / outlet, fp32"


class HipFusedF32S(HipFused):
    """The two-buffer kernel on fp32 storage of f - w_i: the same 4 bytes per population and
    the same speed as ``fused``, with the stored values about w_i / |f - w_i| times finer,
    so steady states driven by small forces keep their accuracy (no reference
    counterpart; the adjoint rollouts store their states this way)."""

    variant = "fused-f32s"
    description = "hand-written HIP (gfx950), fused pull, two buffers, fp32 storage of f - w_i"
    storage = "f32s"


class HipFusedAA(HipFused):
    variant = "fused-aa"
    description = "hand-written HIP (gfx950), AA-pattern in-place streaming, fp32"
    aa = True


class HipFusedFP16(HipFused):
    variant = "fused-fp16"
    description = "hand-written HIP (gfx950), fp16 storage of f - w_i, fp32 arithmetic"
    storage = "f16"


class HipFusedBF16(HipFused):
    variant = "fused-bf16"
    description = "hand-written HIP (gfx950), bf16 storage of f - w_i, fp32 arithmetic"
    storage = "bf16"


class HipFusedAAFP16(HipFused):
    variant = "fused-aa-fp16"
    description = "hand-written HIP (gfx950), AA pattern + fp16 storage of f - w_i"
    storage = "f16"
    aa = True


class HipFusedAABF16(HipFused):
    variant = "fused-aa-bf16"
    description = "hand-written HIP (gfx950), AA pattern + bf16 storage of f - w_i"
    storage = "bf16"
    aa = True


class HipFused2Step(HipFused):
    """Two time steps per launch on D2Q9 boxes (periodic axes or bounce-back walls, the
    lid): the first step's populations stay in LDS, so the state crosses memory once per
    two steps (no reference counterpart; the values are those of two ``fused`` steps)."""

    variant = "fused-2step"
    description = ("hand-written HIP (gfx950), fused pull, two steps per launch, state between "
                   "them on chip, fp32")
    pair = True
    lattices = ("D2Q9",)

    @property
    def bytes_per_node(self) -> int:
        """Minimal traffic of one node update: the state is read once and written once per
        two steps, so q populations cross memory per update."""
        return self.problem.lattice.q * self.itemsize

    def memory_bytes(self) -> int:
        return 2 * self.problem.lattice.q * self.itemsize * self.problem.n_nodes


class HipFused2StepFP16(HipFused2Step):
    variant = "fused-2step-fp16"
    description = ("hand-written HIP (gfx950), fused pull, two steps per launch, fp16 storage "
                   "of f - w_i")
    storage = "f16"


class HipFused2StepBF16(HipFused2Step):
    variant = "fused-2step-bf16"
    description = ("hand-written HIP (gfx950), fused pull, two steps per launch, bf16 storage "
                   "of f - w_i")
    storage = "bf16"
