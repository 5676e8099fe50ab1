"""General geometry on the GPU: solid nodes, velocity inlet, pressure outlet (k_step_geo).

Mirrors the reference's general-geometry equivalence test (tests/test_equivalence.py:78-92:
the Schaefer-Turek channel at d = 6, 60 steps, within 3e-6 of the NumPy solver) with the
golden vectors of tests/golden/geometry.npz, adds random tiny geometries against the
float64 oracle, checks the device link mask bit for bit against np.roll, checks that an
all-fluid solid field reproduces the box kernel bit for bit at 4096^2, and runs the
reference's Karman Strouhal check (tests/test_validation.py:53-58).
"""

from __future__ import annotations

import zlib

import numpy as np
import pytest

from oracle import lbm_ref as R
from pylbm_amd import backends as B
from pylbm_amd.backends import Unsupported
from pylbm_amd.problem import Problem
from pylbm_amd.validation import karman
from tests.cases import GEO_TRAJ, TOL, golden, karman6_kwargs

pytestmark = pytest.mark.gpu

AB_VARIANTS = ["gather", "fused", "fused-f32s", "fused-fp16", "fused-bf16"]

#: fp32 bound: the reference's (tests/test_equivalence.py:21).
TOL_F32 = 3e-6


def rounding_bound(case: R.Case, steps: int, storage: str, f_ref: np.ndarray) -> float:
    """Bound for a 16-bit run: 4x the deviation from the float64 trajectory of the oracle
    re-rounded to the storage format after every step (initial state included), never
    below the reference's tolerance.  The reference's tolerances hold for its box cases,
    where most of the domain is nearly at rest; these channels flow everywhere at u ~ U, so
    |f - w| and its rounding are larger, and at tau near 1/2 (tau = 0.509 for the d = 6
    channel) TRT relaxes the odd part at s- = 0.046 and the rounding accumulates almost
    undamped."""
    lat = case.lat

    def rnd(f):
        return R.decode(R.encode(f, storage, False, lat), storage, False, lat)

    f = rnd(R.initial_populations(case))
    for _ in range(steps):
        f = rnd(R.step(f, case))
    return max(4.0 * float(np.abs(f - f_ref).max()), TOL[storage])


def bound(case: R.Case, steps: int, storage: str, f_ref: np.ndarray) -> float:
    if storage in ("f32", "f32s"):
        return TOL_F32
    return rounding_bound(case, steps, storage, f_ref)


@pytest.fixture(scope="module")
def geo():
    return golden("geometry.npz")


def _make(variant: str, p: Problem):
    cls = B.get("hip", variant)
    ok, why = cls.available()
    assert ok, f"hip/{variant} must be available where a GPU is present: {why}"
    return cls(p)


@pytest.mark.parametrize("variant", AB_VARIANTS)
@pytest.mark.parametrize("coll", ["trt", "bgk", "mrt"])
def test_karman6_matches_reference(variant, coll, geo):
    s = _make(variant, Problem(**karman6_kwargs(geo, coll)))
    s.step(60)
    ref = geo[f"karman6_{coll}/f"]
    err = np.abs(s.populations() - ref).max()
    tol = bound(R.Case(**karman6_kwargs(geo, coll)), 60, s.storage, ref)
    print(f"karman6 {coll} hip/{variant}: max|df| = {err:.3e} (bound {tol:.2e})")
    assert err < tol


@pytest.mark.parametrize("variant", AB_VARIANTS)
@pytest.mark.parametrize("name", list(GEO_TRAJ))
def test_obstacle_trajectory_matches_reference(variant, name, geo):
    s = _make(variant, Problem(**GEO_TRAJ[name], solid=geo[f"{name}/solid"]))
    s.step(60)
    ref = geo[f"{name}/f"]
    err = np.abs(s.populations() - ref).max()
    tol = bound(R.Case(**GEO_TRAJ[name], solid=geo[f"{name}/solid"]), 60, s.storage, ref)
    print(f"{name} hip/{variant}: max|df| = {err:.3e} (bound {tol:.2e})")
    assert err < tol


RANDOM_GEO = {
    "walls_lid_inlet_outlet": dict(shape=(37, 23), u_lid=0.08, inlet=True, outlet=True,
                                   collision="trt"),
    "perx_outlet_force": dict(shape=(41, 19), periodic=(True, False), outlet=True,
                              force=(2e-5, 1e-6), collision="bgk"),
    "pery_inlet_mrt": dict(shape=(33, 17), periodic=(False, True), inlet=True, outlet=True,
                           collision="mrt"),
    "multi_segment_rows": dict(shape=(300, 9), inlet=True, outlet=True, collision="trt"),
    # rows long enough for 2 (fp32) and 4 (16-bit) nodes per lane, with a partial segment
    "wide_rows_lid": dict(shape=(1100, 7), u_lid=0.05, inlet=True, outlet=True,
                          collision="mrt"),
    "wide_rows_periodic_x": dict(shape=(1030, 6), periodic=(True, False), outlet=True,
                                 force=(1e-5, 0.0), collision="bgk"),
    "d3q19_lid": dict(shape=(19, 11, 9), periodic=(True, False, False), u_lid=0.05,
                      collision="trt", force=(1e-5, 0.0, 0.0)),
    "d3q19_closed": dict(shape=(13, 12, 10), collision="bgk"),
}


@pytest.mark.parametrize("name", list(RANDOM_GEO))
def test_random_geometry_matches_oracle(name):
    kw = dict(RANDOM_GEO[name])
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    shape = kw["shape"]
    ashape = tuple(reversed(shape))
    kw["solid"] = rng.random(ashape) < 0.15
    if kw.pop("inlet", False):
        kw["inlet"] = 0.06 * rng.random(shape[1])
    kw.setdefault("tau", 0.65)
    case = R.Case(**kw)
    p = Problem(**kw)
    ref = R.run(case, 40)
    for variant in AB_VARIANTS:
        s = _make(variant, p)
        s.step(40)
        err = np.abs(s.populations() - ref).max()
        tol = 1e-6 if s.storage in ("f32", "f32s") else rounding_bound(case, 40, s.storage,
                                                                        ref)
        print(f"{name} hip/{variant}: max|df| = {err:.3e} (bound {tol:.2e})")
        assert err < tol


@pytest.mark.parametrize("shape", [(29, 7), (6, 5, 4)])
def test_link_mask_equals_rolled_solid(shape):
    rng = np.random.default_rng(11)
    ashape = tuple(reversed(shape))
    solid = rng.random(ashape) < 0.3
    s = _make("gather", Problem(shape, periodic=(True,) * len(shape), solid=solid))
    got = s.links.cpu().numpy().astype(np.int64) & (2 ** s.q - 1)
    lat = s.problem.lattice
    want = np.zeros(ashape, dtype=np.int64)
    for i in range(1, lat.q):
        sh = tuple(int(lat.c[i, a]) for a in reversed(range(len(shape))))
        want |= np.roll(solid, sh, axis=tuple(range(len(shape)))).astype(np.int64) << i
    assert np.array_equal(got.reshape(ashape), want)


@pytest.mark.parametrize("shape,periodic,node", [((29, 7), (True, True), (5, 3)),
                                                 ((130, 5), (False, True), (90, 2)),
                                                 ((200, 3, 2), (False, True, True), (100, 1, 0)),
                                                 ((6, 5, 4), (True, True, True), (2, 2, 1))])
def test_link_groups_flag_the_groups_that_need_bounce_back(shape, periodic, node):
    """pylbm_hip_link_groups: one word per 64-node group of a row, nonzero exactly when a
    node of the group has a solid pull source or sits on a wall of a non-periodic x (one
    solid node: some groups flagged, others not)."""
    ashape = tuple(reversed(shape))
    solid = np.zeros(ashape, bool)
    solid[tuple(reversed(node))] = True
    s = _make("gather", Problem(shape, periodic=periodic, solid=solid))
    nx = shape[0]
    mask = s.links.cpu().numpy().astype(np.int64).reshape(-1, nx) & (2 ** s.q - 1)
    gpr = (nx + 63) // 64
    want = np.zeros((mask.shape[0], gpr), dtype=bool)
    for j in range(gpr):
        x0, x1 = 64 * j, min(64 * j + 64, nx)
        want[:, j] = (mask[:, x0:x1] != 0).any(axis=1)
        if not periodic[0] and (x0 == 0 or x1 == nx):
            want[:, j] = True
    got = s.link_groups.cpu().numpy().reshape(-1, gpr)
    assert set(np.unique(got)) <= {0, 1}
    assert np.array_equal(got != 0, want)
    assert want.any() and not want.all()


GROUP_CASES = {
    "2d-walls-inlet-outlet": (dict(shape=(300, 40), inlet=True, outlet=True, collision="trt"),
                              0.01),
    # rows of 1100: four nodes per lane at 16-bit, a partial segment, the lid
    "2d-wide-rows-lid": (dict(shape=(1100, 24), u_lid=0.05, collision="mrt"), 0.002),
    "2d-periodic-force": (dict(shape=(1030, 16), periodic=(True, True), force=(1e-5, 0.0),
                               collision="bgk"), 0.003),
    "3d-spheres": (dict(shape=(96, 40, 24), periodic=(True, True, True),
                        force=(1e-5, 0.0, 0.0), collision="trt"), None),
    # walls on y and z (x periodic: an x wall would flag every group of its rows)
    "3d-walls-lid": (dict(shape=(140, 12, 10), periodic=(True, False, False), u_lid=0.04,
                          collision="bgk"), 0.001),
}


@pytest.mark.parametrize("variant", ["gather", "fused-f32s", "fused-fp16", "fused-bf16"])
@pytest.mark.parametrize("name", list(GROUP_CASES))
def test_group_flags_give_the_same_state(name, variant):
    """With the group flags (the backend's default) the geometry kernel loads the bounce-back
    values of flagged waves with the pull and skips the mask elsewhere; without them
    (params.link_groups = NULL) it loads them after the mask.  The values and the arithmetic
    are the same: equal states bit for bit."""
    import torch

    kw, frac = GROUP_CASES[name]
    kw = dict(kw)
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    shape = kw["shape"]
    ashape = tuple(reversed(shape))
    if frac is None:  # balls of radius 6 on a pitch of 16 nodes
        c = [(np.arange(n) % 16) - 7.5 for n in ashape]
        kw["solid"] = (c[0][:, None, None] ** 2 + c[1][None, :, None] ** 2
                       + c[2][None, None, :] ** 2) <= 36.0
    else:
        kw["solid"] = rng.random(ashape) < frac
    if kw.pop("inlet", False):
        kw["inlet"] = 0.06 * rng.random(shape[1])
    kw.setdefault("tau", 0.62)
    p = Problem(**kw)
    a, b = _make(variant, p), _make(variant, p)
    assert a.params.link_groups and b.params.link_groups
    groups = a.link_groups.cpu().numpy()
    assert groups.any() and not groups.all()  # both paths of the kernel run
    b.params.link_groups = None
    a.step(30)
    b.step(30)
    torch.cuda.synchronize()
    assert a.reduce("maxabs_diff", b) == 0.0
    assert torch.equal(a.populations_device(), b.populations_device())


@pytest.mark.parametrize("variant", ["fused", "fused-fp16"])
def test_all_fluid_geometry_kernel_equals_box_kernel_4096(variant):
    """Full headline size: the geometry kernel with an all-fluid solid field gives the box
    kernel's state bit for bit (same arithmetic, same loads)."""
    n = 4096
    p0 = Problem((n, n), tau=0.6, periodic=(False, False), u_lid=0.05, collision="trt")
    p1 = Problem((n, n), tau=0.6, periodic=(False, False), u_lid=0.05, collision="trt",
                 solid=np.zeros((n, n), bool))
    a, b = _make(variant, p0), _make(variant, p1)
    assert b.params.links and not a.params.links
    a.step(20)
    b.step(20)
    assert a.reduce("maxabs_diff", b) == 0.0


def test_closed_box_with_obstacles_conserves_fluid_mass_2048():
    import torch

    n = 2048
    rng = np.random.default_rng(5)
    solid = rng.random((n, n)) < 0.1
    s = _make("gather", Problem((n, n), tau=0.7, u_lid=0.05, collision="trt", solid=solid))
    fluid = torch.as_tensor(~solid, device=s.device)
    m0 = float(s.populations_device()[:, fluid].sum())
    s.step(200)
    m1 = float(s.populations_device()[:, fluid].sum())
    print(f"fluid mass {m0:.6f} -> {m1:.6f}")
    assert abs(m1 - m0) < 1e-5 * m0


@pytest.mark.parametrize("variant", ["fused-aa", "fused-aa-fp16", "fused-aa-bf16"])
def test_aa_variants_reject_geometry(variant):
    p = Problem((16, 8), solid=np.zeros((8, 16), bool))
    ok, why = B.get("hip", variant).supports(p)
    assert not ok and "box" in why
    with pytest.raises(Unsupported):
        B.get("hip", variant)(p)


def test_karman_strouhal():
    """tests/test_validation.py:53-58 of the reference: d = 20, 40 + 12 convective units."""
    r = karman.run_karman(d=20, t_transient=40.0, periods=12.0)
    print(f"Karman d=20: St={r['St']:.4f} cd_max={r['cd_max']:.3f} cl_max={r['cl_max']:.3f}")
    st_lo, st_hi = karman.REFERENCE["St"]
    assert 0.95 * st_lo < r["St"] < 1.05 * st_hi
    assert np.isfinite(r["cd_max"]) and 3.0 < r["cd_max"] < 3.6
