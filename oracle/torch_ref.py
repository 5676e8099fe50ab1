"""Float64 torch restatement of the oracle step for autograd -- TEST INFRASTRUCTURE ONLY.

The same algorithm as oracle/lbm_ref.py (pylbm/core.py:88-266), written with out-of-place
torch ops so torch autograd differentiates a K-step rollout with respect to a per-cell Guo
force field, tau and the initial velocity.  The reference has no gradient path (its torch
rungs run under ``torch.no_grad()``, pylbm/backends/torch_backend.py:74); this restatement
is the float64 cross-check of the HIP adjoint.  Streaming uses the per-link rule (a link
that crosses any wall bounces back the node's own opposite population, plus Ladd's term
through the lid), which tests/test_oracle.py pins against the reference's stream.
"""

from __future__ import annotations

import numpy as np
import torch
from torch.utils.checkpoint import checkpoint

from . import lbm_ref as R


class TorchLBM:
    """Differentiable float64 stepping of one :class:`oracle.lbm_ref.Case` geometry."""

    def __init__(self, case: R.Case, device: str = "cpu") -> None:
        self.case = case
        self.lat = case.lat
        self.dev = device
        d, q = case.dim, self.lat.q
        self.c = torch.tensor(self.lat.c, dtype=torch.float64, device=device)
        self.w = torch.tensor(self.lat.w, dtype=torch.float64, device=device)
        self.shifts, self.wall, self.lid = [], [], []
        for i in range(q):
            self.shifts.append(tuple(int(self.lat.c[i, a]) for a in reversed(range(d))))
            _, wall, top = R._neigh(case, i, -1)
            self.wall.append(torch.tensor(wall, device=device))
            lid = np.where(wall & top, 6.0 * self.lat.w[i] * self.lat.c[i, 0] * case.u_lid, 0.0)
            self.lid.append(torch.tensor(lid, dtype=torch.float64, device=device))

    # ------------------------------------------------------------------ pieces
    def equilibrium(self, rho, u):
        usq = sum(u[a] * u[a] for a in range(self.case.dim))
        out = []
        for i in range(self.lat.q):
            cu = sum(float(self.lat.c[i, a]) * u[a] for a in range(self.case.dim))
            out.append(self.lat.w[i] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq))
        return torch.stack(out)

    def stream(self, f):
        d, q = self.case.dim, self.lat.q
        dims = tuple(range(d))
        out = []
        for i in range(q):
            g = torch.roll(f[i], self.shifts[i], dims) if any(self.shifts[i]) else f[i]
            if bool(self.wall[i].any()):
                g = torch.where(self.wall[i], f[self.lat.opp[i]] + self.lid[i], g)
            out.append(g)
        return out

    def rates(self, tau):
        if self.case.collision == "bgk":
            return (1.0 / tau,)
        if self.case.collision == "trt":
            sp = 1.0 / tau
            return (sp, 1.0 / (0.5 + self.case.magic / (1.0 / sp - 0.5)))
        raise NotImplementedError("the adjoint cross-check covers BGK and TRT")

    def collide(self, f, F, tau):
        lat, d, q = self.lat, self.case.dim, self.lat.q
        rho = f[0]
        for i in range(1, q):
            rho = rho + f[i]
        u = []
        for a in range(d):
            j = sum(float(lat.c[i, a]) * f[i] for i in range(q) if lat.c[i, a] != 0)
            u.append((j + 0.5 * F[a]) / rho)
        usq = sum(u[a] * u[a] for a in range(d))
        uf = sum(u[a] * F[a] for a in range(d))
        feq, src = [], []
        for i in range(q):
            cu = sum(float(lat.c[i, a]) * u[a] for a in range(d))
            feq.append(lat.w[i] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq))
            cf = sum(float(lat.c[i, a]) * F[a] for a in range(d))
            src.append(lat.w[i] * (3.0 * (cf - uf) + 9.0 * cu * cf))
        r = self.rates(tau)
        out = []
        if len(r) == 1:
            om = r[0]
            for i in range(q):
                out.append(f[i] - om * (f[i] - feq[i]) + (1.0 - 0.5 * om) * src[i])
        else:
            sp, sm = r
            for i in range(q):
                o = lat.opp[i]
                out.append(f[i] - sp * (0.5 * (f[i] + f[o]) - 0.5 * (feq[i] + feq[o]))
                           - sm * (0.5 * (f[i] - f[o]) - 0.5 * (feq[i] - feq[o]))
                           + (1.0 - 0.5 * sp) * 0.5 * (src[i] + src[o])
                           + (1.0 - 0.5 * sm) * 0.5 * (src[i] - src[o]))
        return torch.stack(out)

    def step(self, f, F, tau):
        return self.collide(self.stream(f), F, tau)

    def velocity(self, f, F):
        """Post-collision u = (sum c f - F/2) / rho."""
        lat, d, q = self.lat, self.case.dim, self.lat.q
        rho = f.sum(0)
        return torch.stack([(sum(float(lat.c[i, a]) * f[i] for i in range(q)
                                 if lat.c[i, a] != 0) - 0.5 * F[a]) / rho for a in range(d)])

    def rollout(self, f0, F, tau, steps: int, segment: int = 0):
        """f after ``steps`` steps; ``segment`` > 0 checkpoints every ``segment`` steps."""
        f = f0
        if segment <= 0:
            for _ in range(steps):
                f = self.step(f, F, tau)
            return f

        def seg(fi, Fi, ti, n=segment):
            for _ in range(n):
                fi = self.step(fi, Fi, ti)
            return fi

        done = 0
        while done < steps:
            n = min(segment, steps - done)
            f = checkpoint(seg, f, F, tau, n, use_reentrant=False)
            done += n
        return f
