"""Differentiable K-step rollouts on the GPU: checkpointed forward + fused HIP adjoint.

:func:`rollout` advances a box problem ``steps`` steps with the two-buffer fused kernels
(fp32, fp16 or bf16 storage) and returns the final post-collision velocity
u = (sum c f - F/2) / rho.  :func:`rollout` wraps it in ``torch.autograd.Function`` so any
torch loss of u back-propagates to

* the per-cell Guo force field F (d, *array_shape),
* tau (BGK omega = 1/tau; TRT s+ = 1/tau, s- from the magic parameter), and
* the initial velocity (and density) of the equilibrium start state.

Forward states are checkpointed in HBM every ``checkpoint`` steps; the backward pass
recomputes one segment at a time, then runs the adjoint steps of that segment in reverse
(``pylbm_hip_adjoint_run``).  What the recompute keeps for the adjoint depends on the
gradients asked for:

* tau not differentiated (the default ``records=None``): every step's moment record, the
  density and velocity its collision used (16 bytes per node,
  ``pylbm_hip_run_record``), while the state itself ping-pongs between two buffers; the
  adjoint reads the records (``pylbm_hip_adjoint_run_record``);
* tau differentiated: every state of the segment (2 q or 4 q bytes per node), since the
  rate sensitivities need the pulled populations.

The adjoint states are fp32 by default; ``lam_storage="f16"`` stores them as fp16 (all
adjoint arithmetic stays fp32), which cuts an adjoint step's traffic from q*(s + 8) to
q*(s + 4) bytes per node at a gradient error of a few 1e-4 relative (measured against
float64 in the tests).  The last segment's states (or records) are kept from the forward
pass, so it is not recomputed.  The reference has no gradient path; the gradients are
checked against float64 torch autograd of a restated step and against central finite
differences.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Any

import torch

from . import _lib
from .problem import Problem

_TORCH_DTYPES = {"f32": torch.float32, "f32s": torch.float32, "f16": torch.float16,
                 "bf16": torch.bfloat16}
_LAM_DTYPES = {"f32": torch.float32, "f16": torch.float16}

#: fp16 adjoint states: the terminal gradient is scaled by a power of two so that its
#: largest component lands near this value, well inside fp16's normal range
#: [6.1e-5, 65504] for the states it seeds (|lam| <= 2 |dL/du| / rho at the seed, and the
#: adjoint of the diffusive step does not grow it); every gradient is divided by the scale.
_LAM16_TARGET = 2.0 ** 8


def lam16_scale(gmax: float) -> float:
    """Power-of-two scale for fp16 adjoint states seeded by a gradient of max-abs gmax."""
    if not (gmax > 0.0) or gmax != gmax or gmax == float("inf"):
        return 1.0
    return 2.0 ** round(math.log2(_LAM16_TARGET / gmax))


def _p(t: Any) -> int | None:
    return None if t is None else int(t.data_ptr())


def _s() -> int:
    return int(torch.cuda.current_stream().cuda_stream)


def tau_rate_derivatives(problem: Problem, tau: float) -> tuple[float, float]:
    """(ds+/dtau, ds-/dtau) of the rates the kernels use (BGK: (d omega/dtau, 0))."""
    dsp = -1.0 / tau ** 2
    if problem.collision == "bgk":
        return dsp, 0.0
    lam = problem.magic
    tau_m = 0.5 + lam / (tau - 0.5)
    return dsp, lam / ((tau - 0.5) ** 2 * tau_m ** 2)


@dataclass
class RolloutConfig:
    """Geometry, collision and storage of a differentiable rollout."""

    problem: Problem  # box, collision bgk|trt; its force_field marks a force-field rollout
    steps: int
    checkpoint: int = 100
    #: "f32s" (default) stores f - w in fp32: the same 4 bytes as "f32", but the
    #: non-equilibrium part that tau acts on keeps its digits; "f16"/"bf16" halve the bytes
    storage: str = "f32s"
    #: element type of the adjoint states between steps: "f32" or "f16" (fp32 arithmetic)
    lam_storage: str = "f32"
    #: recompute the next checkpoint segment on a second stream during each segment's
    #: adjoint (False: one stream, one after the other; the gradients are the same).
    #: None (default): overlap unless the adjoint reads moment records on a grid of 2^20
    #: nodes or more.  On small grids neither kernel fills the GPU alone (1024 x 256
    #: channel: 25.4% -> 28.8% of the forward rate); on large ones both are bandwidth bound
    #: and run concurrently they evict each other's working set from the Infinity Cache
    #: (128^3 box with records: 60.0 ms side by side, 55.6 ms one after the
    #: other; with stored states 59.0 against 60.3 ms)
    overlap: bool | None = None
    #: what the adjoint reads of each forward step: None (default) moment records when tau
    #: is not differentiated, else the stored states; False always the stored states; True
    #: always records (then tau must not be differentiated).  A record is the density and
    #: velocity the step's collision used, 16 bytes per node, which is all the adjoint
    #: needs unless it also sums the rate sensitivities; the stored state is 2 q bytes
    #: (16-bit) or 4 q bytes per node, and the recompute of a segment then keeps its state
    #: in two buffers instead of one per step.
    records: bool | None = None

    def __post_init__(self) -> None:
        p = self.problem
        if self.lam_storage not in _LAM_DTYPES:
            raise ValueError(f"lam_storage must be one of {sorted(_LAM_DTYPES)}")
        if p.collision not in ("bgk", "trt"):
            raise ValueError("the adjoint covers BGK and TRT")
        if not p.is_box():
            raise ValueError("box problems only")
        if self.steps < 1 or self.checkpoint < 1:
            raise ValueError("steps and checkpoint must be >= 1")


class _Engine:
    """Device buffers and C-ABI calls of one rollout configuration."""

    def __init__(self, cfg: RolloutConfig, force_field: torch.Tensor | None,
                 tau: float) -> None:
        self.cfg = cfg
        p = cfg.problem
        self.p = p
        self.q, self.d, self.N = p.lattice.q, p.dim, p.n_nodes
        self.dev = torch.device("cuda", torch.cuda.current_device())
        self.dtype = _TORCH_DTYPES[cfg.storage]
        self.tau = float(tau)
        self.ff = force_field
        prm = _lib.Params()
        prm.lattice = _lib.LATTICE_IDS[p.lattice.name]
        prm.collision = _lib.COLLISION_IDS[p.collision]
        prm.storage = _lib.STORAGE_IDS[cfg.storage]
        prm.aa = 0
        per = tuple(p.periodic) + (True,) * (3 - p.dim)
        shp = tuple(p.shape) + (1,) * (3 - p.dim)
        for k in range(3):
            prm.periodic[k] = int(per[k])
        prm.nx, prm.ny, prm.nz = shp
        prm.n_nodes = self.N
        prm.ulid = float(p.u_lid)
        force = tuple(p.force) + (0.0,) * (3 - p.dim)
        for k in range(3):
            prm.force[k] = force[k]
        rates = Problem(p.shape, tau=self.tau, collision=p.collision, magic=p.magic,
                        periodic=p.periodic).runtime_rates()
        for k, r in enumerate(rates):
            prm.rates[k] = r
        prm.force_field = _p(force_field)
        prm.lam_storage = _lib.STORAGE_IDS[cfg.lam_storage]
        self.prm = prm
        self.ref = ctypes.byref(prm)

    def state(self) -> torch.Tensor:
        return torch.empty(self.q * self.N, device=self.dev, dtype=self.dtype)

    def record(self) -> torch.Tensor:
        """Moment records of one step: (rho, u_x, u_y, u_z) fp32 per node."""
        return torch.empty(4 * self.N, device=self.dev, dtype=torch.float32)

    def init(self, rho0: torch.Tensor | None, u0: torch.Tensor | None, f: torch.Tensor) -> None:
        _lib.call("pylbm_hip_init_equilibrium", self.ref, _p(rho0), _p(u0), _p(f), _s())

    def step(self, src: torch.Tensor, dst: torch.Tensor) -> None:
        _lib.call("pylbm_hip_step", self.ref, _p(src), _p(dst), _lib.AB, _s())

    def velocity(self, f: torch.Tensor) -> torch.Tensor:
        u = torch.empty((self.d,) + self.p.array_shape, dtype=torch.float64, device=self.dev)
        _lib.call("pylbm_hip_macroscopic", self.ref, _p(f), None, _p(u), _s())
        return u


def _ptr_array(tensors: list) -> Any:
    """Host array of device pointers (None -> NULL)."""
    arr = (ctypes.c_void_p * max(len(tensors), 1))()
    for k, t in enumerate(tensors):
        arr[k] = None if t is None else t.data_ptr()
    return arr


class _RolloutFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, force_field, tau, u0, rho0, cfg):  # noqa: ANN001
        ff = None if force_field is None else force_field.detach().to(
            torch.float32).contiguous()
        eng = _Engine(cfg, ff, float(tau))
        u0d = None if u0 is None else u0.detach().to(torch.float64).contiguous()
        r0d = None if rho0 is None else rho0.detach().to(torch.float64).contiguous()
        T, C = cfg.steps, cfg.checkpoint
        K = (T + C - 1) // C
        last0 = (K - 1) * C  # first step of the last segment
        need_tau = bool(ctx.needs_input_grad[1])
        records = (not need_tau) if cfg.records is None else bool(cfg.records)
        if records and need_tau:
            raise ValueError("records=True: the tau gradient needs the stored states")
        f0 = eng.state()
        eng.init(r0d, u0d, f0)
        # where each f^{t+1} goes: its own buffer if it is a checkpoint or f^T (or, without
        # records, an input of the last segment), else one of two scratch buffers (never
        # the buffer it is read from); with records, the steps of the last segment write
        # their records
        scratch = [eng.state(), eng.state()]
        ckpts, last, dst, rec = [f0], [], [], []
        cur = f0
        for t in range(T):
            if t >= last0:
                if records:
                    rec.append(eng.record())
                    last.append(rec[-1])
                else:
                    last.append(cur)
            else:
                rec.append(None)
            keep = (t + 1 == T) or ((t + 1) % C == 0) or (not records and t + 1 >= last0)
            nxt = eng.state() if keep else (scratch[0] if cur is not scratch[0] else scratch[1])
            if t + 1 < T and (t + 1) % C == 0:
                ckpts.append(nxt)
            dst.append(nxt)
            cur = nxt
        if records:
            _lib.call("pylbm_hip_run_record", eng.ref, _p(f0), _ptr_array(dst),
                      _ptr_array(rec), T, _s())
        else:
            _lib.call("pylbm_hip_run_store", eng.ref, _p(f0), _ptr_array(dst), T, _s())
        del scratch, dst, rec
        fT = cur
        u = eng.velocity(fT)
        ctx.eng, ctx.ckpts, ctx.last, ctx.fT = eng, ckpts, last, fT
        ctx.records = records
        ctx.u0d, ctx.r0d = u0d, r0d
        ctx.needs = (force_field is not None and ctx.needs_input_grad[0], need_tau,
                     u0 is not None and ctx.needs_input_grad[2],
                     rho0 is not None and ctx.needs_input_grad[3])
        return u

    @staticmethod
    def backward(ctx, grad_u):  # noqa: ANN001
        eng: _Engine = ctx.eng
        cfg = eng.cfg
        T, C = cfg.steps, cfg.checkpoint
        K = (T + C - 1) // C
        q, d, N = eng.q, eng.d, eng.N
        G = grad_u.detach().to(torch.float32).contiguous()
        scale = 1.0
        if cfg.lam_storage == "f16":
            scale = lam16_scale(float(G.abs().max()))
            G = G * scale
        ldt = _LAM_DTYPES[cfg.lam_storage]
        lam = [torch.empty(q * N, dtype=ldt, device=eng.dev) for _ in range(2)]
        dF = torch.zeros(d * N, dtype=torch.float32, device=eng.dev) if eng.ff is not None \
            else None
        need_F, need_tau, need_u, need_rho = ctx.needs
        # the rate sensitivities cost the kernels a per-population sum and a workgroup
        # reduction: only when dL/dtau is wanted
        part = None
        if need_tau:
            nw = int(_lib.load().pylbm_hip_adjoint_workspace(eng.ref))
            part = torch.zeros(max(nw, 2), dtype=torch.float64, device=eng.dev)
        _lib.call("pylbm_hip_adjoint_seed", eng.ref, _p(ctx.fT), _p(G), _p(lam[0]), _p(dF),
                  _s())
        cur = 0  # index of the lam buffer holding the current adjoint state
        which = ctypes.c_int(0)
        # Segment k is recomputed from checkpoint k into buffer set k % 2 on a second stream
        # (default priority) while the main stream runs the adjoint of segment k + 1: the two
        # are independent, and on small grids neither fills the GPU alone.  Every adjoint
        # step (and so every dF / partial-sum accumulation) stays on the main stream in
        # reverse order, so the gradients do not depend on the overlap.
        records = ctx.records
        overlap = cfg.overlap
        if overlap is None:
            overlap = not records or eng.N < (1 << 20)
        main = torch.cuda.current_stream(eng.dev)
        side = torch.cuda.Stream(device=eng.dev) if overlap else main
        side.wait_stream(main)  # the checkpoints and the seed
        # per buffer set: the segment's stored states (n - 1 of them: the checkpoint is the
        # first), or, with records, two scratch states and the n steps' records
        seg_sets: list[list[torch.Tensor]] = [[], []]
        seg_recs: list[list[torch.Tensor]] = [[], []]
        ready: dict[int, torch.cuda.Event] = {}
        done: dict[int, torch.cuda.Event] = {}

        def recompute(k: int) -> None:
            n = min(C, T - k * C)
            buf, recs = seg_sets[k % 2], seg_recs[k % 2]
            while len(buf) < (2 if records else n - 1):
                buf.append(eng.state())
            while records and len(recs) < n:
                recs.append(eng.record())
            if k + 2 in done:  # the adjoint of segment k + 2 read this buffer set
                side.wait_event(done[k + 2])
            with torch.cuda.stream(side):
                if records:
                    _lib.call("pylbm_hip_run_record", eng.ref, _p(ctx.ckpts[k]),
                              _ptr_array([buf[j % 2] for j in range(n)]), _ptr_array(recs[:n]),
                              n, int(side.cuda_stream))
                else:
                    _lib.call("pylbm_hip_run_store", eng.ref, _p(ctx.ckpts[k]),
                              _ptr_array(buf[:n - 1]), n - 1, int(side.cuda_stream))
            ev = torch.cuda.Event()
            ev.record(side)
            ready[k] = ev

        for k in range(K - 1, -1, -1):
            n = min(C, T - k * C)
            if k == K - 1:
                inputs = ctx.last
            else:
                main.wait_event(ready.pop(k))
                inputs = seg_recs[k % 2][:n] if records else \
                    [ctx.ckpts[k]] + seg_sets[k % 2][:n - 1]
            if k >= 1:  # enqueued before this segment's adjoint, so the two overlap
                recompute(k - 1)
            if records:
                _lib.call("pylbm_hip_adjoint_run_record", eng.ref, _ptr_array(inputs), n,
                          _p(lam[cur]), _p(lam[1 - cur]), _p(dF), ctypes.byref(which),
                          int(main.cuda_stream))
            else:
                _lib.call("pylbm_hip_adjoint_run", eng.ref, _ptr_array(inputs), n,
                          _p(lam[cur]), _p(lam[1 - cur]), _p(dF), _p(part),
                          ctypes.byref(which), int(main.cuda_stream))
            cur = cur if which.value == 0 else 1 - cur
            ev = torch.cuda.Event()
            ev.record(main)
            done[k] = ev
            if k == K - 1:
                ctx.last = None
        main.wait_stream(side)  # the segment buffers are freed after the side stream's work
        lam0 = lam[cur]
        du0 = drho0 = None
        if need_u or need_rho:
            du = torch.empty((d,) + eng.p.array_shape, dtype=torch.float64, device=eng.dev) \
                if need_u else None
            dr = torch.empty(eng.p.array_shape, dtype=torch.float64, device=eng.dev) \
                if need_rho else None
            _lib.call("pylbm_hip_adjoint_init", eng.ref, _p(ctx.r0d), _p(ctx.u0d), _p(lam0),
                      _p(du), _p(dr), _s())
            du0 = du if du is None or scale == 1.0 else du / scale
            drho0 = dr if dr is None or scale == 1.0 else dr / scale
        dtau = None
        if need_tau:
            sums = part.view(-1, 2).sum(0)
            a, b = tau_rate_derivatives(eng.p, eng.tau)
            dtau = ((sums[0] * a + sums[1] * b) / scale).reshape(())
        dFo = dF.view((d,) + eng.p.array_shape) if (dF is not None and need_F) else None
        if dFo is not None and scale != 1.0:
            dFo = dFo / scale
        ctx.eng = ctx.ckpts = ctx.fT = None
        return dFo, dtau, du0, drho0, None


def rollout(cfg: RolloutConfig, force_field: torch.Tensor | None = None,
            tau: torch.Tensor | float | None = None, u0: torch.Tensor | None = None,
            rho0: torch.Tensor | None = None) -> torch.Tensor:
    """Final post-collision velocity (d, *array_shape), float64, differentiable.

    ``force_field`` (d, *array_shape) must be given exactly when ``cfg.problem`` has a
    force field; ``tau`` defaults to ``cfg.problem.tau``; ``u0`` / ``rho0`` default to rest.
    """
    p = cfg.problem
    if (force_field is None) != (p.force_field is None):
        raise ValueError("pass force_field exactly when the problem has a force_field")
    if tau is None:
        tau = torch.tensor(p.tau, dtype=torch.float64, device="cuda")
    elif not torch.is_tensor(tau):
        tau = torch.tensor(float(tau), dtype=torch.float64, device="cuda")
    if float(tau.detach()) <= 0.5:
        raise ValueError("tau must be > 0.5")
    return _RolloutFn.apply(force_field, tau, u0, rho0, cfg)


def velocity_loss(u: torch.Tensor, u_target: torch.Tensor) -> torch.Tensor:
    """0.5 * sum over cells of |u - u_target|^2."""
    return 0.5 * ((u - u_target) ** 2).sum()
