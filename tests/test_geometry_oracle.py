"""The oracle's general geometry (solid nodes, velocity inlet, pressure outlet) against the
reference's own NumPy solver (golden vectors in tests/golden/geometry.npz), a brute-force
per-node rule and conservation; plus the host side of the hip geometry path (problem
checks, the Schaefer-Turek problem builder, force masks).  CPU only."""

from __future__ import annotations

import itertools

import numpy as np
import pytest

from oracle import lbm_ref as R
from pylbm_amd.problem import Problem
from pylbm_amd.validation import karman
from tests.cases import GEO_STREAM, GEO_TRAJ, geo_stream_kwargs, golden, karman6_kwargs


@pytest.fixture(scope="module")
def geo():
    return golden("geometry.npz")


@pytest.mark.parametrize("name", list(GEO_STREAM))
def test_geometry_stream_equals_reference(geo, name):
    case = R.Case(**geo_stream_kwargs(geo, name))
    out = R.stream(geo[f"{name}/in"], case)
    assert np.array_equal(out, geo[f"{name}/out"])


def _brute_stream(f, case):
    """Per node and direction, straight from the rule: the source x - c_i (wrapped on every
    axis) being solid or lying behind a wall bounces back the own opp(i) value; the additive
    term comes from the LAST walled axis crossed (the lid if that is the top wall of the
    last axis, the inlet profile if it is the x = 0 wall); the outlet then sets the c_x = -1
    populations of the last column to the rho = 1 equilibrium of the column to its left."""
    lat, d = case.lat, case.dim
    out = np.empty_like(f)
    solid = np.zeros(case.array_shape, bool) if case.solid is None else case.solid
    for i in range(lat.q):
        for idx in itertools.product(*[range(s) for s in case.array_shape]):
            x = [idx[d - 1 - a] for a in range(d)]
            src, last = [], None
            for a in range(d):
                s = x[a] - lat.c[i, a]
                if (s < 0 or s >= case.shape[a]) and not case.periodic[a]:
                    last = (a, "high" if s >= case.shape[a] else "low")
                src.append(s % case.shape[a])
            sidx = tuple(src[d - 1 - k] for k in range(d))
            if last is not None or (i > 0 and solid[sidx]):
                v = f[(lat.opp[i],) + idx]
                if last == (d - 1, "high") and case.u_lid != 0.0:
                    v = v + 6.0 * lat.w[i] * lat.c[i, 0] * case.u_lid
                if last == (0, "low") and case.inlet is not None:
                    v = v + 6.0 * lat.w[i] * lat.c[i, 0] * case.inlet[x[1]]
            else:
                v = f[(i,) + sidx]
            out[(i,) + idx] = v
    if case.outlet:
        for y in range(case.shape[1]):
            col = f[:, y, -2]
            rho = sum(col[k] for k in range(lat.q))
            ux = sum(lat.c[k, 0] * col[k] for k in range(lat.q)) / rho
            uy = sum(lat.c[k, 1] * col[k] for k in range(lat.q)) / rho
            for i in range(lat.q):
                if lat.c[i, 0] == -1:
                    cu = -ux + lat.c[i, 1] * uy
                    out[i, y, -1] = lat.w[i] * (1 + 3 * cu + 4.5 * cu * cu
                                                - 1.5 * (ux * ux + uy * uy))
    return out


@pytest.mark.parametrize("name", list(GEO_STREAM))
def test_geometry_stream_equals_brute_force_rule(geo, name):
    case = R.Case(**geo_stream_kwargs(geo, name))
    f = geo[f"{name}/in"]
    assert np.allclose(R.stream(f, case), _brute_stream(f, case), rtol=0, atol=1e-15)


@pytest.mark.parametrize("coll", ["trt", "bgk", "mrt"])
def test_karman6_trajectory_equals_reference(geo, coll):
    f = R.run(R.Case(**karman6_kwargs(geo, coll)), 60)
    assert np.abs(f - geo[f"karman6_{coll}/f"]).max() <= 1e-14


@pytest.mark.parametrize("name", list(GEO_TRAJ))
def test_obstacle_trajectory_equals_reference(geo, name):
    f = R.run(R.Case(**GEO_TRAJ[name], solid=geo[f"{name}/solid"]), 60)
    assert np.abs(f - geo[f"{name}/f"]).max() <= 1e-14


def test_closed_box_with_obstacles_conserves_mass():
    rng = np.random.default_rng(3)
    solid = rng.random((9, 11)) < 0.2
    c = R.Case((11, 9), tau=0.7, u_lid=0.05, collision="trt", solid=solid)
    f0 = R.initial_populations(c) * (1 + 0.01 * rng.standard_normal((9, 9, 11)))
    f = R.run(c, 40, f0)
    # solid nodes are streamed too, but nothing flows from them into the fluid: every fluid
    # population either reaches a fluid node or is reflected, so the fluid mass is conserved
    fluid = ~solid
    m0, m1 = f0[:, fluid].sum(), f[:, fluid].sum()
    assert abs(m1 - m0) < 1e-12 * m0


def test_all_fluid_solid_field_is_the_box():
    c0 = R.Case((10, 7), tau=0.7, u_lid=0.1, collision="trt")
    c1 = R.Case((10, 7), tau=0.7, u_lid=0.1, collision="trt", solid=np.zeros((7, 10), bool))
    assert np.array_equal(R.run(c0, 20), R.run(c1, 20))


def test_schafer_turek_problem_matches_reference_inputs(geo):
    p = karman.schafer_turek_problem(d=6, u_mean=0.05, collision="trt")
    assert p.shape == tuple(int(v) for v in geo["karman6/shape"])
    assert p.tau == float(geo["karman6/tau"])
    assert np.array_equal(p.solid, geo["karman6/solid"])
    assert np.array_equal(p.inlet, geo["karman6/inlet"])
    assert np.array_equal(p.init_u, geo["karman6/init_u"])
    assert p.outlet and p.periodic == (False, False)


def test_force_masks_are_fluid_solid_links():
    p = karman.schafer_turek_problem(d=6)
    m = karman.force_masks(p)
    lat = p.lattice
    ny, nx = p.array_shape
    for i in range(1, lat.q):
        for y, x in zip(*np.nonzero(m[i])):
            assert not p.solid[y, x]
            assert p.solid[(y + lat.c[i, 1]) % ny, (x + lat.c[i, 0]) % nx]
    assert not m[0].any()
    # every solid node next to fluid is reached by some link
    assert m.sum() > 0


def test_strouhal_of_a_pure_sine():
    d, u, every = 20, 0.05, 10
    st = 0.3
    t = np.arange(2000) * every
    cl = np.sin(2 * np.pi * st * u / d * t)
    assert abs(karman.strouhal(cl, every, d, u) - st) < 1e-3


@pytest.mark.parametrize("kw,msg", [
    (dict(shape=(6, 4), solid=np.zeros((6, 4), bool)), "solid must have shape"),
    (dict(shape=(6, 4, 3), inlet=np.zeros(4)), "2D"),
    (dict(shape=(6, 4), inlet=np.zeros(6)), "inlet must have shape"),
    (dict(shape=(6, 4, 3), outlet=True), "2D"),
])
def test_problem_rejects_bad_geometry(kw, msg):
    with pytest.raises(ValueError, match=msg):
        Problem(**kw)
