"""Alternating sweep directions (GPU): multi-step loops give single-launch results bit for bit.

pylbm_hip_run, pylbm_hip_run_store and pylbm_hip_adjoint_run alternate the order in which
workgroups take their work items (every launch for the AA pattern, every step for states
resident in the Infinity Cache and for the geometry kernel).  Every node is computed from
the same inputs whatever the order, so a loop must equal the same launches made one at a
time in the forward order,
bit for bit: this checks that the reversed order still visits every work item exactly
once, including packed short rows and grids whose workgroup count is not a multiple of 8
(where the whole order is reversed instead of each XCD's sequence).
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from pylbm_amd import _lib
from pylbm_amd import backends as B
from pylbm_amd.problem import Problem

pytestmark = pytest.mark.gpu

#: variant, problem keywords; comments give the workgroup count and the order policy
CASES = {
    # 4096^2: 65,536 workgroups, larger than the Infinity Cache: only AA alternates
    "aa-4096-f32": ("fused-aa", dict(shape=(4096, 4096), tau=0.6, periodic=(True, True))),
    "aa-4096-fp16": ("fused-aa-fp16", dict(shape=(4096, 4096), tau=0.6, collision="trt",
                                           periodic=(True, True))),
    "ab-4096-fp16": ("fused-fp16", dict(shape=(4096, 4096), tau=0.6, collision="mrt",
                                        periodic=(True, True))),
    # resident states: 1024 workgroups (two segments per row), slabs, alternating
    "ab-resident-512": ("fused", dict(shape=(512, 512), tau=0.7, collision="trt",
                                      u_lid=0.05)),
    # short rows packed two to a workgroup: 167 workgroups, the whole order reversed
    "ab-short-rows": ("fused", dict(shape=(100, 333), tau=0.65, collision="mrt",
                                    periodic=(True, False), force=(1e-5, 0.0))),
    # D3Q19 64^3, rows packed four to a workgroup: 1024 workgroups
    "ab-resident-3d": ("fused-fp16", dict(shape=(64, 64, 64), tau=0.6, collision="trt",
                                          periodic=(True, True, True))),
    "aa-resident-3d": ("fused-aa-bf16", dict(shape=(96, 40, 24), tau=0.7, collision="trt",
                                             u_lid=0.04)),
    # geometry kernel (alternates at every size): 4 nodes per lane at 16-bit, 2 segments
    "geometry-fp16": ("fused-fp16", dict(shape=(2048, 300), tau=0.6, collision="trt",
                                         periodic=(False, False), outlet=True,
                                         inlet=np.full(300, 0.03))),
    "geometry-3d": ("gather", dict(shape=(40, 24, 16), tau=0.7, collision="bgk",
                                   periodic=(True, False, True), force=(1e-5, 0.0, 0.0))),
}


def _spheres(shape):
    """A solid field of a few balls (discs in 2D) away from the walls."""
    grids = np.meshgrid(*[np.arange(n) for n in reversed(shape)], indexing="ij")
    solid = np.zeros(tuple(reversed(shape)), bool)
    for k in range(3):
        c = [(0.3 + 0.2 * k) * n for n in reversed(shape)]
        solid |= sum((g - ci) ** 2 for g, ci in zip(grids, c)) <= (0.08 * min(shape)) ** 2
    return solid


@pytest.mark.parametrize("name", list(CASES))
def test_run_loop_equals_single_launches(name):
    variant, kw = CASES[name]
    rng = np.random.default_rng(3)
    ashape = tuple(reversed(kw["shape"]))
    u0 = 0.02 * rng.standard_normal((len(ashape),) + ashape)
    n = 6
    if name.startswith("geometry"):
        kw = dict(kw, solid=_spheres(kw["shape"]))
    a, b = (B.make("hip", variant, Problem(**kw), init="none") for _ in range(2))
    for s in (a, b):
        s.init_equilibrium(None, u0)
        if s.B is not None:  # an unwritten slot of a destination then shows as NaN
            s.B.fill_(float("nan"))
    a.step(n)  # one pylbm_hip_run call: alternating directions where the policy says so
    for _ in range(n // 2 if b.aa else n):  # single launches, forward order
        if b.aa:
            b.launch("aa-odd")
            b.launch("aa-even")
        else:
            b.launch("ab")
    d = a.reduce("maxabs_diff", b)  # NaN if either state holds NaN
    print(f"{name}: run loop vs single launches max|df| = {d}")
    assert d == 0.0
    assert torch.equal(a.A, b.A)


ADJ_CASES = [((24, 16), (True, False), 5, 0.0, "f32"),
             ((24, 16), (True, False), 6, 0.0, "f32"),
             ((40, 28), (False, False), 4, 0.0, "f32"),
             ((300, 40), (False, True), 4, 0.0, "f32"),
             ((20, 12, 10), (False, True, False), 6, 0.0, "f32"),
             ((32, 32, 32), (True, True, True), 5, 0.0, "f32"),
             ((32, 32, 32), (True, True, True), 6, 0.0, "f32"),
             ((300, 40), (True, True), 4, 0.0, "f32"),
             # moving lid (its constant term drops out of the adjoint) and fp16 adjoint states
             ((40, 28), (False, False), 6, 0.05, "f32"),
             ((40, 28), (False, False), 5, 0.05, "f16"),
             ((300, 40), (False, True), 4, 0.0, "f16"),
             ((20, 12, 10), (False, True, False), 6, 0.04, "f16"),
             ((32, 32, 32), (True, True, True), 5, 0.0, "f16")]


@pytest.mark.parametrize("shape,periodic,n,u_lid,lam_storage", ADJ_CASES)
def test_adjoint_run_equals_single_adjoint_steps(shape, periodic, n, u_lid, lam_storage):
    """pylbm_hip_adjoint_run (alternating on these resident states; the in-place AA
    pattern in pairs, an odd last step two-buffer; walls on x, y, z and none, with and
    without a lid; fp32 and fp16 adjoint states) vs single two-buffer adjoint steps: the
    same adjoint state and force gradient bit for bit; the dL/ds partial sums land in other
    per-workgroup slots, so their total agrees to float64 rounding."""
    from pylbm_amd.adjoint import _LAM_DTYPES, RolloutConfig, _Engine, _ptr_array

    rng = np.random.default_rng(8)
    d = len(shape)
    ashape = tuple(reversed(shape))
    ff = (1e-5 * rng.standard_normal((d,) + ashape)).astype(np.float32)
    p = Problem(shape, tau=0.7, periodic=periodic, collision="trt", u_lid=u_lid,
                force_field=ff.astype(np.float64))
    q, N = p.lattice.q, p.n_nodes
    ffd = torch.as_tensor(ff, device="cuda").reshape(d, -1)
    eng = _Engine(RolloutConfig(p, 1, storage="f32s", lam_storage=lam_storage), ffd, 0.7)
    states = []
    f = eng.state()
    u0 = torch.as_tensor(0.02 * rng.standard_normal((d,) + ashape), device="cuda")
    eng.init(None, u0, f)
    for _ in range(n):
        states.append(f)
        g = eng.state()
        eng.step(f, g)
        f = g
    lam0 = torch.as_tensor(0.01 * rng.standard_normal(q * N), dtype=torch.float32,
                           device="cuda").to(_LAM_DTYPES[lam_storage])
    nw = int(_lib.load().pylbm_hip_adjoint_workspace(eng.ref))
    out = []
    for mode in ("run", "steps"):
        lam = [lam0.clone(), torch.empty_like(lam0)]
        dF = torch.zeros(d * N, dtype=torch.float32, device="cuda")
        part = torch.zeros(nw, dtype=torch.float64, device="cuda")
        if mode == "run":
            which = ctypes.c_int(0)
            _lib.call("pylbm_hip_adjoint_run", eng.ref, _ptr_array(states), n,
                      lam[0].data_ptr(), lam[1].data_ptr(), dF.data_ptr(), part.data_ptr(),
                      ctypes.byref(which), None)
            res = lam[which.value]
        else:
            cur = 0
            for k in range(n - 1, -1, -1):
                _lib.call("pylbm_hip_adjoint_step", eng.ref, states[k].data_ptr(),
                          lam[cur].data_ptr(), lam[1 - cur].data_ptr(), dF.data_ptr(),
                          part.data_ptr(), None)
                cur = 1 - cur
            res = lam[cur]
        torch.cuda.synchronize()
        out.append((res.cpu(), dF.cpu(), part.view(-1, 2).sum(0).cpu().numpy()))
    assert torch.equal(out[0][0], out[1][0])
    assert torch.equal(out[0][1], out[1][1])
    assert np.allclose(out[0][2], out[1][2], rtol=1e-12, atol=0.0)
