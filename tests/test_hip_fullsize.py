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


@pytest.mark.parametrize("storage", ["", "-fp16"])
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
