"""Every stored slot is written by every launch, and the float64 reductions are exact (GPU).

* Two-buffer steps (pylbm/backends/cuda/lbm.cu:165-198): the destination is filled with
  NaN before each launch; afterwards every slot is finite and equals the state computed
  into a zero-filled destination, bit for bit.  A kernel that skipped a work item, a row
  segment, a node slot of a lane or a population would leave NaN behind (with a
  ``torch.empty`` destination the caching allocator can instead return a buffer holding an
  earlier valid state and hide a missing store).  The C loops (pylbm_hip_run_store, with
  the reversed sweeps of resident states) are covered the same way, one fresh NaN buffer
  per step.  The two-steps-per-launch kernel is covered the same way, per launch.
* AA steps work in place, so their destinations cannot be poisoned.  Instead each launch
  runs on X and on X + eps g, g a constant ghost field that changes no gathered moment:
  with BGK at tau = 1 the two outputs agree to rounding, while a slot the launch did not
  write keeps the offset eps g_i, ten times that rounding (the premise is checked on the
  float64 oracle in tests/test_oracle.py).
* Reductions (mass, kinetic energy, loss, dot, max|f - g|) give the same float64 bits on
  every call, including grids large enough to take the grid-stride path; MAXABS_DIFF
  reports NaN when either state holds a NaN or an Inf - Inf difference.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from pylbm_amd import _lib
from pylbm_amd import backends as B
from pylbm_amd.problem import Problem

pytestmark = pytest.mark.gpu

NAN = float("nan")

SMALL = {
    # walls on both axes, lid corners, 2 row segments with a partial last one
    "2d-walls-lid": dict(shape=(300, 40), tau=0.7, collision="trt", periodic=(False, False),
                         u_lid=0.05),
    "2d-periodic-mrt": dict(shape=(512, 48), tau=0.6, collision="mrt", periodic=(True, True)),
    # rows of 50 nodes packed four to a workgroup, walls in y, uniform force
    "2d-short-rows": dict(shape=(50, 33), tau=0.8, collision="bgk", periodic=(True, False),
                          force=(1e-5, 0.0)),
    # rows of 777: four nodes per lane at 16-bit storage, a partial segment
    "2d-wide-rows": dict(shape=(777, 9), tau=0.65, collision="trt", periodic=(False, True)),
    "3d-walls-lid": dict(shape=(20, 12, 10), tau=0.7, collision="bgk",
                         periodic=(False, True, False), u_lid=0.05),
    "3d-periodic": dict(shape=(128, 16, 8), tau=0.55, collision="trt", periodic=(True,) * 3),
}
TWO_BUFFER = ["fused", "fused-f32s", "fused-fp16", "fused-bf16"]
AA = ["fused-aa", "fused-aa-fp16", "fused-aa-bf16"]
PAIR = ["fused-2step", "fused-2step-fp16", "fused-2step-bf16"]
PAIR_CASES = {
    "2d-periodic-mrt": SMALL["2d-periodic-mrt"],
    # a partial last tile of columns, chunks of 9 rows with a shorter last one, a force
    "2d-partial-tile": dict(shape=(600, 71), tau=0.62, collision="bgk", periodic=(True, True),
                            force=(1e-5, 0.0)),
    # two chunks of rows, the last one row high
    "2d-one-row-chunk": dict(shape=(300, 10), tau=0.7, collision="trt", periodic=(True, True)),
    # walls on both axes and the lid
    "2d-walls-lid": SMALL["2d-walls-lid"],
}


def _geometry_case() -> dict:
    nx, ny = 600, 64
    y, x = np.mgrid[0:ny, 0:nx]
    solid = (x - 150) ** 2 + (y - 30) ** 2 <= 12 ** 2
    return dict(shape=(nx, ny), tau=0.6, collision="trt", periodic=(False, False),
                solid=solid, inlet=np.full(ny, 0.04), outlet=True)


def _random_state(p: Problem, seed: int, amp: float = 5e-3) -> np.ndarray:
    """A canonical float64 state far from equilibrium: equilibrium of a random field plus
    random non-equilibrium parts on every population."""
    from pylbm_amd.lattice import equilibrium

    rng = np.random.default_rng(seed)
    ashape = p.array_shape
    rho = 1.0 + 0.01 * rng.standard_normal(ashape)
    u = 0.03 * rng.standard_normal((p.dim,) + ashape)
    f = equilibrium(p.lattice, rho, u)
    w = np.asarray(p.lattice.w).reshape((-1,) + (1,) * p.dim)
    return f + amp * w * rng.standard_normal(f.shape)


def _pair(variant: str, kw: dict, seed: int = 1):
    p = Problem(**kw)
    f0 = _random_state(p, seed)
    out = []
    for _ in range(2):
        s = B.make("hip", variant, p, init="none")
        s.set_populations(f0)
        out.append(s)
    return out


@pytest.mark.parametrize("variant", TWO_BUFFER)
@pytest.mark.parametrize("name", [*SMALL, "geometry"])
def test_two_buffer_step_writes_every_slot(name, variant):
    kw = _geometry_case() if name == "geometry" else SMALL[name]
    poisoned, clean = _pair(variant, kw)
    for k in range(3):
        poisoned.B.fill_(NAN)
        clean.B.zero_()
        poisoned.launch("ab")
        clean.launch("ab")
        torch.cuda.synchronize()
        n_bad = int((~torch.isfinite(poisoned.A.float())).sum())
        assert n_bad == 0, f"{name} hip/{variant} step {k}: {n_bad} slots left unwritten"
        assert torch.equal(poisoned.A, clean.A), f"{name} hip/{variant} step {k}"


@pytest.mark.parametrize("variant", PAIR)
@pytest.mark.parametrize("name", list(PAIR_CASES))
def test_two_step_launch_writes_every_slot(name, variant):
    """Two steps per launch (the intermediate state on chip): every destination slot is
    written once, with the value the launch computes into a zero-filled destination."""
    poisoned, clean = _pair(variant, PAIR_CASES[name])
    for k in range(3):
        poisoned.B.fill_(NAN)
        clean.B.zero_()
        poisoned.launch("ab2")
        clean.launch("ab2")
        torch.cuda.synchronize()
        n_bad = int((~torch.isfinite(poisoned.A.float())).sum())
        assert n_bad == 0, f"{name} hip/{variant} launch {k}: {n_bad} slots left unwritten"
        assert torch.equal(poisoned.A, clean.A), f"{name} hip/{variant} launch {k}"


@pytest.mark.parametrize("name", ["2d-periodic-mrt", "2d-short-rows", "3d-periodic",
                                  "geometry"])
def test_run_store_writes_every_slot_of_every_destination(name):
    """pylbm_hip_run_store over fresh NaN buffers (odd steps reverse the sweep of these
    resident states and of the geometry kernel) equals single forward launches."""
    from pylbm_amd.adjoint import _ptr_array

    kw = _geometry_case() if name == "geometry" else SMALL[name]
    a, b = _pair("fused-fp16", kw)
    dst = [torch.full_like(a.A, NAN) for _ in range(4)]
    _lib.call("pylbm_hip_run_store", a.pref, a.A.data_ptr(), _ptr_array(dst), len(dst),
              torch.cuda.current_stream().cuda_stream)
    for k in range(4):
        b.B.zero_()
        b.launch("ab")
        torch.cuda.synchronize()
        assert bool(torch.isfinite(dst[k].float()).all()), f"{name} step {k}"
        assert torch.equal(dst[k], b.A), f"{name} step {k}"


def _ghost(lat) -> np.ndarray:
    """A constant population field g with sum_i g_i = 0, sum_i c_i g_i = 0, g_opp(i) = g_i and
    every g_i nonzero (g_i depends on |c_i|^2 only; for D2Q9 it is the energy row of the
    MRT basis)."""
    c2 = (lat.c ** 2).sum(axis=1)
    vals = {9: (-4.0, -1.0, 2.0), 19: (30.0, 11.0, -8.0)}[lat.q]
    g = np.choose(c2, vals)
    assert g.sum() == 0 and not (lat.c.T @ g).any() and (g != 0).all()
    assert (g[lat.opp] == g).all()
    return g


#: largest |out(X) - out(X + eps g)| allowed per storage: rounding only (fp32 arithmetic,
#: then the storage's rounding of values of size ~0.05)
GHOST_TOL = {"f32": 1e-6, "f16": 3e-4, "bf16": 2e-3}


@pytest.mark.parametrize("variant", AA)
@pytest.mark.parametrize("name", list(SMALL))
def test_aa_launches_write_every_slot(name, variant):
    """With BGK at tau = 1 a node's output depends on the density and velocity it gathers
    only (p = feq(rho, u) + S/2 with Guo's S(u, F)).  Adding eps g to every stored plane
    (bounce-back reads the own opp(i) slot and g_opp(i) = g_i) changes no gathered moment, so
    each AA launch must map X and X + eps g to the same array up to rounding; a slot the
    launch did not write keeps its offset eps g_i, 10x above that rounding."""
    kw = dict(SMALL[name], tau=1.0, collision="bgk")
    kw.pop("force", None)
    if name == "2d-short-rows":
        kw["force"] = (1e-5, 0.0)
    p = Problem(**kw)
    g = _ghost(p.lattice)
    eps = 0.025 if p.lattice.q == 9 else 0.003
    f0 = _random_state(p, 2)
    a, b = (B.make("hip", variant, p, init="none") for _ in range(2))
    for s in (a, b):
        s.set_populations(f0)
    tol = GHOST_TOL[a.storage]
    assert 10 * tol < eps * np.abs(g).min()
    q = p.lattice.q
    gt = torch.as_tensor(eps * g, dtype=torch.float32, device="cuda").view(q, 1)
    for k in range(2):
        for pattern in ("aa-odd", "aa-even"):
            bv = b.A.view(q, -1)
            bv.copy_((bv.float() + gt).to(bv.dtype))
            a.launch(pattern)
            b.launch(pattern)
            torch.cuda.synchronize()
            d = float((a.A.float() - b.A.float()).abs().max())
            assert d < tol, f"{name} hip/{variant} pair {k} {pattern}: {d:.3e}"


# ------------------------------------------------------------------- benchmark sizes


def _equal_planes(x, y, q) -> bool:
    n = x.numel() // q
    return all(torch.equal(x[i * n:(i + 1) * n], y[i * n:(i + 1) * n]) for i in range(q))


def _finite_planes(x, q) -> bool:
    n = x.numel() // q
    return all(bool(torch.isfinite(x[i * n:(i + 1) * n].float()).all()) for i in range(q))


@pytest.mark.parametrize("variant", TWO_BUFFER)
def test_two_buffer_step_writes_every_slot_4096(variant):
    from tests.cases import shear_wave_fields

    rho, u = shear_wave_fields(4096, seed=7)
    p = Problem((4096, 4096), tau=0.6, collision="trt", periodic=(True, True))
    a, b = (B.make("hip", variant, p, init="none") for _ in range(2))
    for s in (a, b):
        s.init_equilibrium(rho, u)
        s.step(3)
    a.B.fill_(NAN)
    b.B.zero_()
    a.launch("ab")
    b.launch("ab")
    torch.cuda.synchronize()
    assert _finite_planes(a.A, 9)
    assert _equal_planes(a.A, b.A, 9)


@pytest.mark.parametrize("variant", ["fused", "fused-fp16", "fused-bf16"])
def test_two_buffer_step_writes_every_slot_512_cubed(variant):
    """D3Q19 512^3 (arrays above 4 GiB: plane-group descriptors): one step into a NaN
    destination against one into a zero-filled destination."""
    n = 512
    p = Problem((n,) * 3, tau=0.55, collision="trt", periodic=(True,) * 3)
    k = 2 * np.pi * torch.arange(n, dtype=torch.float64, device="cuda") / n
    z, y, x = k.view(n, 1, 1), k.view(1, n, 1), k.view(1, 1, n)
    u = torch.zeros((3, n, n, n), dtype=torch.float64, device="cuda")
    u[0] = 0.02 * torch.sin(x) * torch.cos(y) * torch.cos(z)
    u[1] = -0.02 * torch.cos(x) * torch.sin(y) * torch.cos(z)
    a = B.make("hip", variant, p, init="none")
    a.init_equilibrium(None, u)
    a.step(1)
    ref = torch.zeros_like(a.A)
    poisoned = torch.full_like(a.A, NAN)
    del u
    st = torch.cuda.current_stream().cuda_stream
    _lib.call("pylbm_hip_step", a.pref, a.A.data_ptr(), ref.data_ptr(), _lib.AB, st)
    _lib.call("pylbm_hip_step", a.pref, a.A.data_ptr(), poisoned.data_ptr(), _lib.AB, st)
    torch.cuda.synchronize()
    ok_finite = _finite_planes(poisoned, 19)
    ok_equal = _equal_planes(poisoned, ref, 19)
    del a, ref, poisoned
    torch.cuda.empty_cache()
    assert ok_finite and ok_equal


# ------------------------------------------------------------------------- reductions


@pytest.mark.parametrize("shape", [(300, 40), (2048, 2048), (96, 64, 80)])
def test_reductions_are_bitwise_reproducible(shape):
    """Every reduction kind returns the same float64 on repeated calls; the 2048^2 grid
    (16,384 row-workgroups) takes the grid-stride path of the capped grid.  The sums also
    agree with float64 sums on the host to 1e-12."""
    dim = len(shape)
    p = Problem(shape, tau=0.7, periodic=(True,) * dim, force=(1e-5,) + (0.0,) * (dim - 1))
    a, b = _pair("fused", dict(shape=shape, tau=0.7, periodic=(True,) * dim,
                               force=(1e-5,) + (0.0,) * (dim - 1)), seed=4)
    b.step(1)
    rng = np.random.default_rng(9)
    target = 0.01 * rng.standard_normal((dim,) + p.array_shape)
    got = {}
    for kind, other in (("mass", None), ("kinetic", None), ("loss", target), ("dot", b),
                        ("maxabs_diff", b)):
        vals = [a.reduce(kind, other) for _ in range(4)]
        assert len({v.hex() for v in vals}) == 1, f"{shape} {kind}: {vals}"
        got[kind] = vals[0]
    fa, fb = a.populations(), b.populations()
    assert abs(got["mass"] - fa.sum()) <= 1e-12 * abs(fa.sum())
    assert abs(got["dot"] - (fa * fb).sum()) <= 1e-12 * abs((fa * fb).sum())
    assert got["maxabs_diff"] == np.abs(fa - fb).max()


@pytest.mark.parametrize("variant", ["fused", "fused-fp16"])
def test_maxabs_diff_reports_nan_and_inf(variant):
    kw = dict(shape=(300, 40), tau=0.7, periodic=(True, True))
    a, b = _pair(variant, kw)
    assert a.reduce("maxabs_diff", b) == 0.0
    n = a.A.numel()
    a.A[n // 3] = NAN
    assert np.isnan(a.reduce("maxabs_diff", b))
    assert np.isnan(b.reduce("maxabs_diff", a))
    a.A[n // 3] = float("inf")
    assert a.reduce("maxabs_diff", b) == float("inf")
    b.A[n // 3] = float("inf")
    assert np.isnan(a.reduce("maxabs_diff", b))  # inf - inf
    assert a.reduce("mass") == float("inf")
