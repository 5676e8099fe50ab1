"""Two time steps per launch (the fused-2step variants, PYLBM_AB2, pylbm_hip_run_pairs),
on periodic boxes and with walls (half-way bounce-back, the lid).

The kernel keeps the first step's populations in LDS, encoded exactly as a stored state, so
its result is the result of two single two-buffer steps: the same gather, decode and
collision on the same values.  Checked against the single-step kernel (bit for bit),
against the float64 oracle within the reference's tolerances, and at the headline size.  CPU tests cover the supports() and ABI rules.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle import lbm_ref as R
from pylbm_amd import _lib
from pylbm_amd import backends as B
from pylbm_amd.problem import Problem
from tests.cases import TOL

VARIANTS = {"f32": ("fused", "fused-2step"), "f16": ("fused-fp16", "fused-2step-fp16"),
            "bf16": ("fused-bf16", "fused-2step-bf16")}


# ----------------------------------------------------------------------------- CPU


def test_2step_supports_d2q9_boxes_only():
    cls = B.get("hip", "fused-2step")
    assert cls.supports(Problem((64, 32), periodic=(True, True)))[0]
    assert cls.supports(Problem((64, 32), periodic=(True, False)))[0]
    assert cls.supports(Problem((64, 32), u_lid=0.1))[0]
    ok, why = cls.supports(Problem((8, 8, 8), periodic=(True, True, True)))
    assert not ok and "D3Q19" in why
    ok, why = cls.supports(Problem((64, 32), periodic=(True, True),
                                   solid=np.zeros((32, 64), bool)))
    assert not ok and "box" in why
    ok, why = cls.supports(Problem((64, 32), inlet=np.zeros(32), outlet=True))
    assert not ok and "box" in why


def test_abi_rejects_two_steps_where_they_do_not_apply():
    lib = _lib.load()
    p = _lib.Params()
    p.lattice, p.collision, p.storage, p.aa = _lib.D2Q9, _lib.BGK, _lib.F32, 0
    p.periodic[0], p.periodic[1], p.periodic[2] = 1, 0, 1
    p.nx, p.ny, p.nz, p.n_nodes = 16, 8, 1, 128
    p.rates[0] = 1.0
    a, b = ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000)
    # obstacles (a link mask) are not a box problem
    p.links = ctypes.c_void_p(0x3000)
    assert lib.pylbm_hip_step(ctypes.byref(p), a, b, _lib.AB2, None) == 1003
    assert "box" in lib.pylbm_hip_last_error().decode()
    assert lib.pylbm_hip_run_pairs(ctypes.byref(p), a, b, 4, None) == 1003
    p.links = None
    # a D2Q9 state of 4 GiB or more
    p.nx, p.ny, p.n_nodes = 16384, 8192, 16384 * 8192
    assert lib.pylbm_hip_run_pairs(ctypes.byref(p), a, b, 4, None) == 1003
    assert "4 GiB" in lib.pylbm_hip_last_error().decode()
    p.nx, p.ny, p.n_nodes = 16, 8, 128
    assert lib.pylbm_hip_run_pairs(ctypes.byref(p), a, b, 3, None) == 1001
    assert lib.pylbm_hip_run_pairs(ctypes.byref(p), a, a, 4, None) == 1001
    assert lib.pylbm_hip_step(ctypes.byref(p), a, a, _lib.AB2, None) == 1001


# ----------------------------------------------------------------------------- GPU

CASES = {
    # rows of three tiles (the last partial), rows not a multiple of the chunk height
    "600x70-bgk": dict(shape=(600, 70), tau=0.6, collision="bgk"),
    "600x70-trt-force": dict(shape=(600, 70), tau=0.62, collision="trt", force=(1e-5, -2e-6)),
    "600x70-mrt": dict(shape=(600, 70), tau=0.58, collision="mrt"),
    # one partial tile, few rows
    "40x7-trt": dict(shape=(40, 7), tau=0.7, collision="trt"),
    "256x3-bgk": dict(shape=(256, 3), tau=0.8, collision="bgk", force=(2e-5, 1e-5)),
    "1x5-mrt": dict(shape=(1, 5), tau=0.7, collision="mrt"),
    # one row, one node, two rows, rows of 21 tiles (the last partial) in chunks of one row
    "300x1-trt": dict(shape=(300, 1), tau=0.66, collision="trt"),
    "1x1-bgk": dict(shape=(1, 1), tau=0.9, collision="bgk", force=(1e-5, 2e-5)),
    "257x2-mrt": dict(shape=(257, 2), tau=0.6, collision="mrt"),
    "5300x6-bgk": dict(shape=(5300, 6), tau=0.7, collision="bgk"),
    # walls: the lid-driven cavity (corners, a partial tile), a channel (walls in y only),
    # walls in x only, one column, two rows
    "120x70-cavity-bgk": dict(shape=(120, 70), tau=0.6, collision="bgk", u_lid=0.05,
                              periodic=(False, False)),
    "600x40-cavity-trt-force": dict(shape=(600, 40), tau=0.7, collision="trt", u_lid=0.05,
                                    force=(1e-5, 2e-6), periodic=(False, False)),
    "300x50-channel-mrt": dict(shape=(300, 50), tau=0.62, collision="mrt", force=(2e-5, 0.0),
                               periodic=(True, False)),
    "64x33-xwalls-bgk-field": dict(shape=(64, 33), tau=0.65, collision="bgk", field=True,
                                   periodic=(False, True)),
    "1x9-walls-trt": dict(shape=(1, 9), tau=0.8, collision="trt", u_lid=0.1,
                          periodic=(False, False)),
    "257x2-cavity-mrt": dict(shape=(257, 2), tau=0.7, collision="mrt", u_lid=0.03,
                             periodic=(False, False)),
    # a per-cell force field
    "300x40-bgk-field": dict(shape=(300, 40), tau=0.6, collision="bgk", field=True),
    "300x40-mrt-field": dict(shape=(300, 40), tau=0.65, collision="mrt", field=True),
}


def _problem(kw: dict) -> Problem:
    kw = dict(kw)
    kw.setdefault("periodic", (True, True))
    if kw.pop("field", False):
        nx, ny = kw["shape"]
        kw["force_field"] = 1e-5 * np.random.default_rng(5).standard_normal((2, ny, nx))
    return Problem(**kw)


def _state(p, rng):
    shp = p.array_shape
    u = np.zeros((2,) + shp)
    u[0] = 0.03 * np.sin(2 * np.pi * np.arange(shp[0]) / shp[0])[:, None]
    u[1] = 0.01 * rng.standard_normal(shp)
    rho = 1.0 + 1e-3 * rng.uniform(-1, 1, shp)
    return rho, u


def _ulps(a, b, storage):
    """Largest distance of two stored states in units in the last place of the storage:
    the sign-magnitude bit patterns mapped onto ordered integers (-0 and +0 coincide, and a
    sign change counts the steps through zero)."""
    import torch

    if storage == "f32":
        ia, ib, top = a.view(torch.int32).long(), b.view(torch.int32).long(), 1 << 31
    else:
        ia, ib, top = a.view(torch.int16).long(), b.view(torch.int16).long(), 1 << 15

    def ordered(i):
        return torch.where(i < 0, -(i + top), i)

    return int((ordered(ia) - ordered(ib)).abs().max())


@pytest.mark.gpu
@pytest.mark.parametrize("storage", list(VARIANTS))
@pytest.mark.parametrize("name", list(CASES))
def test_2step_equals_two_single_steps(name, storage):
    """10 launches of two steps against 20 single steps from the same state: bit for bit
    (the collisions fuse multiply-adds within a statement only, so both kernels round
    alike)."""
    import torch

    kw = CASES[name]
    rng = np.random.default_rng(21)
    single, double = VARIANTS[storage]
    p = _problem(kw)
    rho, u = _state(p, rng)
    a = B.make("hip", single, p, init="none")
    b = B.make("hip", double, p, init="none")
    a.init_equilibrium(rho, u)
    b.init_equilibrium(rho, u)
    a.step(20)
    b.step(20)
    torch.cuda.synchronize()
    d = _ulps(a.A, b.A, storage)
    print(f"{name} {storage}: largest difference {d} ulp")
    assert d == 0


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["fused-2step", "fused-2step-fp16", "fused-2step-bf16"])
@pytest.mark.parametrize("coll", ["bgk", "trt", "mrt"])
def test_2step_matches_oracle(variant, coll):
    """The periodic shear wave of the benchmark at 256 x 128 (rows of one tile, periodic
    wrap on both axes), 100 steps, against the float64 oracle within the reference's
    tolerances."""
    rng = np.random.default_rng(22)
    kw = dict(shape=(256, 128), tau=0.6, collision=coll, periodic=(True, True))
    if coll == "mrt":
        kw.update(s_e=1.2, s_eps=1.2)
    p = Problem(**kw)
    rho, u = _state(p, rng)
    case = R.Case(**kw)
    f0 = R.equilibrium(case.lat, rho, u)
    ref = R.run(case, 100, f0)
    s = B.make("hip", variant, p, init="none")
    s.init_equilibrium(rho, u)
    s.step(100)
    err = float(np.abs(s.populations() - ref).max())
    tol = TOL[s.storage]
    print(f"{variant} {coll}: max|df| = {err:.2e} (tol {tol:.0e})")
    assert err < tol


@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_2step_at_4096_equals_single_steps(storage):
    """The headline size (16 tiles per row, over a hundred chunks of rows): 20 steps as 10
    launches of two equal 20 single steps bit for bit, and conserve mass."""
    import torch

    n = 4096
    rng = np.random.default_rng(23)
    p = Problem((n, n), tau=0.6, periodic=(True, True))
    rho, u = _state(p, rng)
    single, double = VARIANTS[storage]
    a = B.make("hip", single, p, init="none")
    b = B.make("hip", double, p, init="none")
    a.init_equilibrium(rho, u)
    b.init_equilibrium(rho, u)
    m0 = b.reduce("mass")
    a.step(20)
    b.step(20)
    torch.cuda.synchronize()
    d = _ulps(a.A, b.A, storage)
    m1 = b.reduce("mass")
    print(f"4096^2 {storage}: largest difference {d} ulp; mass {m0:.10e} -> {m1:.10e}")
    assert d == 0
    assert abs(m1 - m0) < (1e-9 if storage == "f32" else 1e-6) * m0


def test_abi_rejects_bad_sweep_probe_calls():
    lib = _lib.load()
    a, b = ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000)
    assert lib.pylbm_hip_sweep_probe(3, a, b, 64, 64, None) == 1001
    assert lib.pylbm_hip_sweep_probe(4, a, a, 64, 64, None) == 1001
    assert lib.pylbm_hip_sweep_probe(4, a, b, 0, 64, None) == 1001
    assert lib.pylbm_hip_sweep_probe(4, a, b, 16384, 16384, None) == 1002
    assert "4 GiB" in lib.pylbm_hip_last_error().decode()
