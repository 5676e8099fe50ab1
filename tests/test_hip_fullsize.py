"""Size-independent properties at the benchmark sizes (GPU): D2Q9 4096^2 and D3Q19 512^3.

The oracle cannot run these sizes in seconds, so the checks are properties that hold
whatever the size:

* the AA pattern gives the two-buffer state (pylbm/codegen.py:10-16) to within a couple of
  units in the last place of the storage format (the two kernels evaluate the same
  collision on the same values; the compiler contracts multiply-adds differently in
  them), for every collision and storage, as the reference requires (< 1e-6 in fp32,
  tests/test_equivalence.py:55-62);
* translation invariance on a periodic box: shifting the initial state by k nodes and
  stepping equals stepping and shifting, bit for bit, for shifts that move nodes across
  workgroup segments, node-per-lane slots and the row / plane wrap;
* mass conservation on a periodic box without force (float64 device reduction).
"""

from __future__ import annotations

import numpy as np
import pytest

from pylbm_amd import backends as B
from pylbm_amd.problem import Problem
from tests.cases import shear_wave_fields

pytestmark = pytest.mark.gpu

N2 = 4096


def _make(variant: str, p: Problem):
    cls = B.get("hip", variant)
    ok, why = cls.available()
    assert ok, f"hip/{variant} must be available where a GPU is present: {why}"
    return cls(p, init="none")


def _maxdiff(a, b) -> float:
    """max |a - b| over two canonical float64 device states, one population at a time."""
    x, y = a.populations_device(), b.populations_device()
    m = max(float((x[i] - y[i]).abs().max()) for i in range(x.shape[0]))
    del x, y
    return m


@pytest.fixture(scope="module")
def shear4096():
    return shear_wave_fields(N2, seed=2026)


def _p2(coll: str) -> Problem:
    extra = {"s_e": 1.2, "s_eps": 1.2} if coll == "mrt" else {}
    return Problem((N2, N2), tau=0.6, collision=coll, periodic=(True, True), **extra)


#: AA vs two-buffer after 20 steps (observed: f32 6.0e-8, fp16 6.0e-6, bf16 4.3e-5)
AA_TOL = {"": 1e-6, "-fp16": 5e-5, "-bf16": 4e-4}


@pytest.mark.parametrize("storage", ["", "-fp16", "-bf16"])
@pytest.mark.parametrize("coll", ["bgk", "trt", "mrt"])
def test_aa_equals_two_buffer_4096(coll, storage, shear4096):
    rho, u = shear4096
    ab, aa = _make("fused" + storage, _p2(coll)), _make("fused-aa" + storage, _p2(coll))
    for s in (ab, aa):
        s.init_equilibrium(rho, u)
        s.step(20)
    d = _maxdiff(ab, aa)
    print(f"4096^2 {coll} AB vs AA{storage}: max|df| = {d:.3e}")
    assert d < AA_TOL[storage]


@pytest.mark.parametrize("variant", ["fused", "fused-fp16", "fused-aa"])
@pytest.mark.parametrize("shift", [(0, 1), (0, 300), (0, 1025), (1, 0), (7, 513)])
def test_translation_invariance_4096(variant, shift, shear4096):
    """roll(step(f0)) == step(roll(f0)) for a shift (dy, dx) of the whole periodic box."""
    import torch

    rho, u = shear4096
    dy, dx = shift
    a, b = _make(variant, _p2("trt")), _make(variant, _p2("trt"))
    a.init_equilibrium(rho, u)
    b.init_equilibrium(np.roll(rho, shift, axis=(0, 1)), np.roll(u, shift, axis=(1, 2)))
    a.step(10)
    b.step(10)
    x, y = a.populations_device(), b.populations_device()
    d = max(float((torch.roll(x[i], shifts=(dy, dx), dims=(0, 1)) - y[i]).abs().max())
            for i in range(9))
    print(f"4096^2 hip/{variant} shift {shift}: max|df| = {d:.3e}")
    assert d == 0.0


@pytest.mark.parametrize("variant", ["fused", "fused-bf16", "fused-aa-fp16"])
def test_mass_conservation_4096(variant, shear4096):
    rho, u = shear4096
    s = _make(variant, _p2("mrt"))
    s.init_equilibrium(rho, u)
    m0 = s.reduce("mass")
    s.step(200)
    m1 = s.reduce("mass")
    print(f"4096^2 hip/{variant}: mass {m0:.9e} -> {m1:.9e} ({(m1 - m0) / m0:.2e})")
    assert abs(m1 - m0) <= 1e-6 * m0


def _tgv512():
    import torch

    n = 512
    k = 2 * np.pi * torch.arange(n, dtype=torch.float64, device="cuda") / n
    z, y, x = k.view(n, 1, 1), k.view(1, n, 1), k.view(1, 1, n)
    u = torch.zeros((3, n, n, n), dtype=torch.float64, device="cuda")
    u[0] = 0.02 * torch.sin(x) * torch.cos(y) * torch.cos(z)
    u[1] = -0.02 * torch.cos(x) * torch.sin(y) * torch.cos(z)
    return u


@pytest.mark.parametrize("storage", ["", "-fp16", "-bf16"])
def test_d3q19_512_aa_matches_two_buffer_and_conserves_mass(storage):
    import torch

    p = Problem((512,) * 3, tau=0.55, collision="trt", periodic=(True,) * 3)
    u = _tgv512()
    ab, aa = _make("fused" + storage, p), _make("fused-aa" + storage, p)
    for s in (ab, aa):
        s.init_equilibrium(None, u)
    del u
    m0 = ab.reduce("mass")
    for s in (ab, aa):
        s.step(10)
    m1 = ab.reduce("mass")
    d = _maxdiff(ab, aa)
    torch.cuda.empty_cache()
    print(f"512^3 TRT{storage}: AB vs AA max|df| = {d:.3e}; mass {m0:.9e} -> {m1:.9e}")
    assert d < AA_TOL[storage]
    assert abs(m1 - m0) <= 1e-6 * m0


# ----------------------------------------------------------------- adjoint at the box size


def test_adjoint_in_place_at_128_cubed_equals_two_buffer_steps():
    """The 128^3 periodic box of the differentiable D3Q19 case: pylbm_hip_adjoint_run runs
    pairs of steps in place (AA pattern) with alternating sweeps; single two-buffer adjoint
    steps give the same adjoint state bit for bit, with fp32 and with fp16 adjoint states
    (the per-node algebra fuses multiply-adds within statements only, so every adjoint
    kernel rounds it identically)."""
    import ctypes

    import torch

    from pylbm_amd import _lib
    from pylbm_amd.adjoint import RolloutConfig, _Engine, _ptr_array
    from pylbm_amd.fields import divergence_free_field

    n, steps = 128, 4
    p = Problem((n,) * 3, tau=0.7, periodic=(True,) * 3)
    u0 = torch.as_tensor(divergence_free_field(n, 8, 0.02, 1), device="cuda")
    for ls, dt in (("f32", torch.float32), ("f16", torch.float16)):
        eng = _Engine(RolloutConfig(p, 1, storage="f16", lam_storage=ls), None, 0.7)
        states, f = [], eng.state()
        eng.init(None, u0, f)
        for _ in range(steps):
            states.append(f)
            g = eng.state()
            eng.step(f, g)
            f = g
        q, N = p.lattice.q, p.n_nodes
        lam0 = (0.01 * torch.randn(q * N, device="cuda", generator=torch.Generator(
            device="cuda").manual_seed(3))).to(dt)
        nw = int(_lib.load().pylbm_hip_adjoint_workspace(eng.ref))
        out = []
        for mode in ("run", "steps"):
            lam = [lam0.clone(), torch.empty_like(lam0)]
            part = torch.zeros(nw, dtype=torch.float64, device="cuda")
            if mode == "run":
                which = ctypes.c_int(0)
                _lib.call("pylbm_hip_adjoint_run", eng.ref, _ptr_array(states), steps,
                          lam[0].data_ptr(), lam[1].data_ptr(), None, part.data_ptr(),
                          ctypes.byref(which), None)
                res = lam[which.value]
            else:
                cur = 0
                for k in range(steps - 1, -1, -1):
                    _lib.call("pylbm_hip_adjoint_step", eng.ref, states[k].data_ptr(),
                              lam[cur].data_ptr(), lam[1 - cur].data_ptr(), None,
                              part.data_ptr(), None)
                    cur = 1 - cur
                res = lam[cur]
            torch.cuda.synchronize()
            out.append(res.clone())
        print(f"128^3 adjoint, {ls} states: in place vs two-buffer equal = "
              f"{bool(torch.equal(out[0], out[1]))}")
        assert torch.equal(out[0], out[1])
        del states, out, lam
        torch.cuda.empty_cache()


def test_box3d_rollout_gradient_fp16_states_at_full_size():
    """The differentiable D3Q19 case at its full size (128^3, T = 500, checkpoints every 50,
    fp16-shifted forward storage): the fp16-state gradient dL/du0 stays within fp16
    rounding of the fp32-state gradient, and both rollouts give the same loss."""
    import torch

    from pylbm_amd.adjoint import RolloutConfig, rollout, velocity_loss
    from pylbm_amd.fields import divergence_free_field

    n, T = 128, 500
    p = Problem((n,) * 3, tau=0.7, periodic=(True,) * 3)
    u0 = torch.as_tensor(divergence_free_field(n, 8, 0.02, 1), device="cuda")
    uo = torch.as_tensor(divergence_free_field(n, 8, 0.02, 2), device="cuda")
    res = {}
    for ls in ("f32", "f16"):
        cfg = RolloutConfig(p, T, checkpoint=50, storage="f16", lam_storage=ls)
        U = u0.clone().requires_grad_(True)
        L = velocity_loss(rollout(cfg, u0=U), uo)
        L.backward()
        res[ls] = (float(L), U.grad.clone())
        del U, L
        torch.cuda.empty_cache()
    g32, g16 = res["f32"][1], res["f16"][1]
    err = float((g16 - g32).abs().max() / g32.abs().max())
    print(f"128^3 T={T}: loss {res['f32'][0]:.9e} / {res['f16'][0]:.9e}; "
          f"du0 fp16-state vs fp32-state rel err {err:.2e}")
    assert res["f32"][0] == res["f16"][0]  # same forward rollout
    assert err < 3e-3  # fp16 rounding of the adjoint states (3e-4 at 32^3, T = 100)
