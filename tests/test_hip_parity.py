"""HIP kernels vs the float64 oracle (GPU).

Mirrors the reference's backend-equivalence tests (tests/test_equivalence.py:34-62): every
``hip`` variant reproduces the reference trajectories of the seven box cases after 100
steps (golden vectors written from the reference NumPy solver), within the reference's
tolerances; the AA pattern equals the two-buffer pattern; plus the config-level checks:
the 128^2 cavity after 1000 steps and the periodic 256^2 shear wave (BGK/TRT/MRT, pull and
AA) within 1e-5 of float64.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import lbm_ref as R
from pylbm_amd import backends as B
from pylbm_amd.problem import Problem
from tests.cases import EQUIVALENCE, TOL, golden, shear_wave_fields

pytestmark = pytest.mark.gpu

VARIANTS = [v for _, v in B.entries()]


def _make(variant: str, p: Problem):
    cls = B.get("hip", variant)
    ok, why = cls.available()
    assert ok, f"hip/{variant} must be available where a GPU is present: {why}"
    if cls.pair:
        # two steps per launch run on periodic D2Q9 boxes only; every other variant runs
        # every case
        ok, why = cls.supports(p)
        if not ok:
            pytest.skip(f"hip/{variant}: {why}")
    return cls(p)


@pytest.fixture(scope="module")
def eq_golden():
    return golden("equivalence.npz")


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("case", list(EQUIVALENCE))
def test_matches_reference_trajectory(variant, case, eq_golden):
    p = Problem(**EQUIVALENCE[case])
    s = _make(variant, p)
    s.step(100)
    f = s.populations()
    err = np.abs(f - eq_golden[case]).max()
    print(f"{case} hip/{variant}: max|df| = {err:.3e}")
    assert err < TOL[s.storage]


@pytest.mark.parametrize("variant", ["fused", "fused-aa"])
def test_cavity128_1000_steps(variant):
    """Config 1: D2Q9 BGK cavity 128^2, tau = 0.884, U = 0.1, 1000 steps, max|df| <= 1e-5."""
    g = golden("cavity128.npz")
    p = Problem.cavity(128, re=100.0, u_lid=0.1)
    s = _make(variant, p)
    s.step(1000)
    f = s.populations()
    err = max(np.abs(f[:, ::4, ::4] - g["sub"]).max(), np.abs(f[:, -1, :] - g["lid_row"]).max(),
              np.abs(f[:, :, 0] - g["left_col"]).max())
    print(f"cavity128 hip/{variant}: max|df| = {err:.3e}")
    assert err <= 1e-5
    # fp32 round-off drifts the total mass by ~1e-5 relative over 1000 steps
    assert np.allclose(f.sum(axis=(1, 2)), g["sum"], rtol=1e-4)


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_aa_pattern_equals_two_buffer(storage):
    """tests/test_equivalence.py:55-62: AA and two-buffer agree (< 1e-6 in fp32)."""
    suffix = "" if storage == "f32" else "-" + {"f16": "fp16", "bf16": "bf16"}[storage]
    p = Problem((64, 40), tau=0.6, u_lid=0.1, collision="trt")
    a, b = _make("fused" + suffix, p), _make("fused-aa" + suffix, p)
    a.step(60)
    b.step(60)
    err = np.abs(a.populations() - b.populations()).max()
    print(f"AA vs AB {storage}: {err:.3e}")
    assert err < {"f32": 1e-6, "f16": 2e-4, "bf16": 2e-3}[storage]


SHEAR_COLL = {"bgk": {}, "trt": {}, "mrt": {"s_e": 1.2, "s_eps": 1.2},
              "mrt-sq1.2": {"s_e": 1.2, "s_eps": 1.2, "s_q": 1.2}}
_SHEAR_REF: dict = {}


def _shear_case(name: str) -> dict:
    n = 256
    rho, u = shear_wave_fields(n, seed=7)
    return dict(shape=(n, n), tau=0.6, collision=name.split("-")[0], periodic=(True, True),
                init_rho=rho, init_u=u, **SHEAR_COLL[name])


def _shear_ref(name: str):
    if name not in _SHEAR_REF:
        case = R.Case(**_shear_case(name))
        f = R.run(case, 200)
        _SHEAR_REF[name] = (f, R.macroscopic(f, case))
    return _SHEAR_REF[name]


@pytest.mark.parametrize("variant", [v for v in VARIANTS if v != "gather"])
@pytest.mark.parametrize("coll", list(SHEAR_COLL))
def test_periodic_shear_wave_256(variant, coll):
    """Config 2 same-run check: periodic 256^2 shear wave (rho = 1 + 1e-3 U(-1, 1),
    u_x = 0.01 sin(2 pi y / 256)), tau(+) = 0.6, MRT with s_e = s_eps = 1.2 and the magic or
    a 1.2 s_q, 200 steps, every variant; populations, density and velocity vs float64."""
    f_ref, (rho_ref, u_ref) = _shear_ref(coll)
    s = _make(variant, Problem(**_shear_case(coll)))
    s.step(200)
    err = np.abs(s.populations() - f_ref).max()
    rho, u = s.macroscopic()
    erho, eu = np.abs(rho - rho_ref).max(), np.abs(u - u_ref).max()
    print(f"shear256 {coll} hip/{variant}: max|df| = {err:.3e}  Linf rho {erho:.2e} "
          f"u {eu:.2e}")
    bound = {"f32": 1e-5, "f32s": 1e-5, "f16": TOL["f16"], "bf16": TOL["bf16"]}[s.storage]
    assert err <= bound
    assert eu <= {"f32": 1e-6, "f32s": 1e-6, "f16": 1e-4, "bf16": 1e-3}[s.storage]


@pytest.mark.parametrize("variant", VARIANTS)
def test_state_roundtrip_and_observables(variant):
    """encode/decode, device init and float64 observables equal the host computations."""
    rng = np.random.default_rng(3)
    # forces exact in fp32, so the float64 host formulas see the value the kernels see
    kw = dict(shape=(37, 19), tau=0.8, periodic=(True, False), force=(2.0**-15, -(2.0**-16)))
    p = Problem(**kw)
    case = R.Case(**kw)
    lat = case.lat
    f = lat.w[:, None, None] * (1.0 + 0.05 * rng.standard_normal((9, 19, 37)))
    s = _make(variant, p)
    s.set_populations(f)
    got = s.populations()
    want = R.decode(R.encode(f, s.storage, s.aa, lat), s.storage, s.aa, lat)
    assert np.array_equal(got, want)
    rho, u = s.macroscopic()
    rho_o, u_o = R.macroscopic(got, case)
    assert np.array_equal(rho, rho_o) and np.array_equal(u, u_o)
    assert abs(s.reduce("mass") - got.sum()) < 1e-9
    # device equilibrium init == host equilibrium + encode
    rho0 = 1.0 + 0.01 * rng.standard_normal((19, 37))
    u0 = 0.02 * rng.standard_normal((2, 19, 37))
    s.init_equilibrium(rho0, u0)
    want = R.decode(R.encode(R.equilibrium(lat, rho0, u0), s.storage, s.aa, lat),
                    s.storage, s.aa, lat)
    assert np.array_equal(s.populations(), want)


@pytest.mark.parametrize("variant", ["fused", "fused-aa-bf16"])
def test_loss_and_dot_reductions(variant):
    """Float64 device reductions LOSS (0.5 sum |u - u_t|^2) and DOT (sum f g over canonical
    values) against the host formulas on the decoded states."""
    rng = np.random.default_rng(5)
    kw = dict(shape=(45, 21), tau=0.7, periodic=(True, False), force=(2.0**-15, 0.0))
    p = Problem(**kw)
    case = R.Case(**kw)
    a, b = _make(variant, p), _make(variant, p)
    a.init_equilibrium(1 + 0.01 * rng.standard_normal((21, 45)),
                       0.02 * rng.standard_normal((2, 21, 45)))
    a.step(6)
    b.step(4)
    fa, fb = a.populations(), b.populations()
    _, ua = R.macroscopic(fa, case)
    ut = 0.01 * rng.standard_normal((2, 21, 45))
    want_loss = 0.5 * ((ua - ut) ** 2).sum()
    assert abs(a.reduce("loss", ut) - want_loss) <= 1e-12 * abs(want_loss)
    want_dot = (fa * fb).sum()
    assert abs(a.reduce("dot", b) - want_dot) <= 1e-12 * abs(want_dot)


def test_per_cell_force_field_matches_oracle():
    """Per-cell Guo force (extension): HIP vs oracle on a channel, 100 steps."""
    rng = np.random.default_rng(11)
    nx, ny = 24, 16
    ff = np.zeros((2, ny, nx))
    ff[0] = 1e-5 + 1e-6 * rng.standard_normal((ny, nx))
    ff[1] = 1e-6 * rng.standard_normal((ny, nx))
    ff = ff.astype(np.float32).astype(np.float64)  # the kernels read the field in fp32
    # D3Q19 TRT: walls in y and z with a lid, a 3-component field
    f3 = (1e-5 * rng.standard_normal((3, 6, 8, 12))).astype(np.float32).astype(np.float64)
    cases = [dict(shape=(nx, ny), tau=0.8, periodic=(True, False), collision=coll,
                  force_field=ff) for coll in ("bgk", "trt", "mrt")]
    cases.append(dict(shape=(12, 8, 6), tau=0.7, periodic=(True, False, False), u_lid=0.03,
                      collision="trt", force_field=f3))
    for kw in cases:
        ref = R.run(R.Case(**kw), 100)
        for variant in ("fused", "fused-aa"):
            s = _make(variant, Problem(**kw))
            s.step(100)
            err = np.abs(s.populations() - ref).max()
            print(f"force field {len(kw['shape'])}D {kw['collision']} hip/{variant}: {err:.3e}")
            assert err < 3e-6


# Rows longer than one workgroup segment, with partial last segments: a lane serves 1, 2 or
# 4 nodes one workgroup width (256) apart depending on lattice, storage and pattern, so
# these widths put walls, the lid corner and the periodic wrap on every node slot of a lane.
WIDE_ROWS = {
    "d2_cavity_1100": dict(shape=(1100, 7), tau=0.7, u_lid=0.1, collision="trt"),
    "d2_perx_1030_force": dict(shape=(1030, 6), tau=0.65, periodic=(True, False),
                               force=(1e-5, 2e-6), collision="mrt"),
    "d2_periodic_777": dict(shape=(777, 5), tau=0.6, periodic=(True, True), collision="bgk"),
    "d3_box_600": dict(shape=(600, 3, 4), tau=0.7, u_lid=0.05, collision="trt"),
}


@pytest.mark.parametrize("variant", [v for v in VARIANTS if v != "gather"])
@pytest.mark.parametrize("case", list(WIDE_ROWS))
def test_wide_rows_match_oracle(variant, case):
    kw = dict(WIDE_ROWS[case])
    rng = np.random.default_rng(7)
    ashape = tuple(reversed(kw["shape"]))
    kw["init_u"] = 0.02 * rng.standard_normal((len(ashape),) + ashape)
    p = Problem(**kw)
    ref = R.run(R.Case(**kw), 40)
    s = _make(variant, p)
    s.step(40)
    err = np.abs(s.populations() - ref).max()
    print(f"{case} hip/{variant}: max|df| = {err:.3e}")
    assert err < (1e-6 if s.storage == "f32" else TOL[s.storage])
