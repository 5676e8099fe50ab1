"""Pins the float64 oracle (CPU).

Each part of oracle/lbm_ref.py is checked against something that does not share its code:
the reference's own golden vectors (tests/golden/, written from the reference NumPy solver),
brute-force per-node rule evaluation, conservation laws and reductions to BGK, and the
AA-pattern restatement against the two-buffer step.
"""

from __future__ import annotations

import itertools

import numpy as np
import pytest

from oracle import lbm_ref as R
from tests.cases import EQUIVALENCE, golden

# ----------------------------------------------------------------------------- golden


@pytest.fixture(scope="module")
def units():
    return golden("units.npz")


@pytest.mark.parametrize("case", list(EQUIVALENCE))
def test_trajectory_equals_reference_solver(case):
    g = golden("equivalence.npz")[case]
    f = R.run(R.Case(**EQUIVALENCE[case]), 100)
    assert np.abs(f - g).max() <= 1e-14


def test_cavity128_1000_steps_equals_reference_solver():
    g = golden("cavity128.npz")
    c = R.Case((128, 128), tau=3.0 * 0.1 * 128 / 100.0 + 0.5, u_lid=0.1)
    assert abs(c.tau - 0.884) < 1e-12
    f = R.run(c, 1000)
    assert np.abs(f[:, ::4, ::4] - g["sub"]).max() <= 1e-14
    assert np.abs(f[:, -1, :] - g["lid_row"]).max() <= 1e-14
    assert np.abs(f[:, :, 0] - g["left_col"]).max() <= 1e-14
    assert np.allclose(f.sum(axis=(1, 2)), g["sum"], rtol=1e-13, atol=0)


STREAM_CASES = {
    "s2_walls_lid": dict(shape=(5, 4), u_lid=0.07),
    "s2_perx_lid": dict(shape=(5, 4), periodic=(True, False), u_lid=0.05),
    "s2_pery": dict(shape=(6, 5), periodic=(False, True)),
    "s2_periodic": dict(shape=(5, 3), periodic=(True, True)),
    "s3_walls_lid": dict(shape=(4, 3, 5), u_lid=0.06),
    "s3_mixed_lid": dict(shape=(4, 5, 3), periodic=(True, False, False), u_lid=0.04),
    "s3_perz": dict(shape=(3, 4, 5), periodic=(False, False, True)),
}


@pytest.mark.parametrize("name", list(STREAM_CASES))
def test_stream_equals_reference(units, name):
    out = R.stream(units[f"{name}/in"], R.Case(**STREAM_CASES[name]))
    assert np.array_equal(out, units[f"{name}/out"])


COLLIDE_CASES = {
    "c2_bgk": dict(shape=(7, 6), tau=0.8),
    "c2_bgk_force": dict(shape=(7, 6), tau=0.8, force=(1e-4, -2e-4)),
    "c2_trt_force": dict(shape=(7, 6), tau=0.6, collision="trt", force=(3e-4, 1e-4)),
    "c2_mrt": dict(shape=(7, 6), tau=0.7, collision="mrt"),
    "c2_mrt_force": dict(shape=(7, 6), tau=0.6, collision="mrt", force=(-1e-4, 2e-4),
                         s_e=1.2, s_eps=1.2),
    "c3_bgk_force": dict(shape=(4, 5, 3), tau=0.7, force=(1e-4, 2e-5, -5e-5)),
    "c3_trt": dict(shape=(4, 5, 3), tau=0.55, collision="trt"),
    "c3_trt_force": dict(shape=(4, 5, 3), tau=0.9, collision="trt", force=(-1e-4, 3e-5, 2e-4)),
    "c2_mrt_sq12": dict(shape=(7, 6), tau=0.6, collision="mrt", force=(2e-4, -1e-4), s_e=1.2,
                        s_eps=1.2, s_q=1.2),
}


@pytest.mark.parametrize("name", list(COLLIDE_CASES))
def test_collide_equals_reference(units, name):
    out = R.collide(units[f"{name}/in"], R.Case(**COLLIDE_CASES[name]))
    assert np.abs(out - units[f"{name}/out"]).max() <= 1e-16


# ----------------------------------------------------------------------------- tables


@pytest.mark.parametrize("lat", [R.D2Q9, R.D3Q19])
def test_lattice_isotropy(lat):
    c, w = lat.c.astype(float), lat.w
    assert np.isclose(w.sum(), 1.0)
    assert np.allclose((w[:, None] * c).sum(0), 0.0)
    assert np.allclose(np.einsum("i,ia,ib->ab", w, c, c), np.eye(lat.d) / 3.0)
    assert (c[lat.opp] == -c).all()
    # fourth-order isotropy of the weights: sum w c_a c_b c_c c_d = (d_ab d_cd + ...)/9
    d = lat.d
    four = np.einsum("i,ia,ib,ic,id->abcd", w, c, c, c, c)
    I = np.eye(d)
    want = (np.einsum("ab,cd->abcd", I, I) + np.einsum("ac,bd->abcd", I, I)
            + np.einsum("ad,bc->abcd", I, I)) / 9.0
    assert np.allclose(four, want)


def test_mrt_basis_orthogonal():
    assert np.allclose(R.MRT_M @ R.MRT_M.T, np.diag(R.MRT_NORM))
    assert np.array_equal(R.MRT_NORM, [9, 36, 36, 6, 12, 6, 12, 4, 4])


@pytest.mark.parametrize("lat", [R.D2Q9, R.D3Q19])
def test_equilibrium_moments(lat):
    rng = np.random.default_rng(0)
    rho = 1 + 0.01 * rng.standard_normal((4, 5))
    u = 0.05 * rng.standard_normal((lat.d, 4, 5))
    feq = R.equilibrium(lat, rho, u)
    assert np.allclose(feq.sum(0), rho)
    assert np.allclose(np.einsum("ia,i...->a...", lat.c.astype(float), feq), rho * u)
    # second moment: sum c_a c_b feq = rho (u_a u_b + delta_ab / 3)
    pi = np.einsum("ia,ib,i...->ab...", lat.c.astype(float), lat.c.astype(float), feq)
    want = rho * (u[:, None] * u[None, :] + np.eye(lat.d)[:, :, None, None] / 3.0)
    assert np.allclose(pi, want)


def test_magic_rate_closed_form():
    s = 1 / 0.6
    sq = R.magic_rate(s)
    assert np.isclose((1 / s - 0.5) * (1 / sq - 0.5), 3.0 / 16.0)
    assert np.isclose(sq, 1 / 2.375)  # tau+ = 0.6
    assert np.isclose(R.magic_rate(1 / 0.55), 1 / 4.25)


# ----------------------------------------------------------------------------- collision


def _rand_f(lat, shape, seed):
    rng = np.random.default_rng(seed)
    return lat.w.reshape((-1,) + (1,) * len(shape)) * (
        1 + 0.05 * rng.standard_normal((lat.q,) + shape))


@pytest.mark.parametrize("force", [None, (1e-4, -2e-4)])
def test_mrt_with_equal_rates_is_bgk(force):
    tau = 0.8
    f = _rand_f(R.D2Q9, (6, 7), 1)
    a = R.collide(f, R.Case((7, 6), tau=tau, force=force))
    b = R.collide(f, R.Case((7, 6), tau=tau, force=force, collision="mrt", s_e=1 / tau,
                            s_eps=1 / tau, s_q=1 / tau))
    assert np.allclose(a, b, atol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_trt_with_equal_rates_is_bgk(dim):
    tau = 0.9
    magic = (tau - 0.5) ** 2  # makes s- = s+
    shape = (5, 4) if dim == 2 else (3, 4, 2)
    lat = R.lattice_for(dim)
    f = _rand_f(lat, tuple(reversed(shape)), 2)
    force = (1e-4,) * dim
    a = R.collide(f, R.Case(shape, tau=tau, force=force))
    b = R.collide(f, R.Case(shape, tau=tau, force=force, collision="trt", magic=magic))
    assert np.allclose(a, b, atol=1e-14)


@pytest.mark.parametrize("coll", ["bgk", "trt", "mrt"])
def test_collision_conserves_mass_and_momentum(coll):
    f = _rand_f(R.D2Q9, (6, 7), 2)
    c = R.Case((7, 6), tau=0.7, collision=coll)
    g = R.collide(f, c)
    cc = R.D2Q9.c.astype(float)
    assert np.allclose(g.sum(0), f.sum(0), atol=1e-15)
    assert np.allclose(np.einsum("ia,iyx->ayx", cc, g), np.einsum("ia,iyx->ayx", cc, f),
                       atol=1e-15)


@pytest.mark.parametrize("coll", ["bgk", "trt", "mrt"])
def test_guo_force_adds_momentum(coll):
    """With force F, collision adds exactly F to the momentum of every node."""
    f = _rand_f(R.D2Q9, (6, 7), 3)
    F = (3e-4, -1e-4)
    g = R.collide(f, R.Case((7, 6), tau=0.7, collision=coll, force=F))
    cc = R.D2Q9.c.astype(float)
    dj = np.einsum("ia,iyx->ayx", cc, g) - np.einsum("ia,iyx->ayx", cc, f)
    assert np.allclose(dj[0], F[0], atol=1e-15) and np.allclose(dj[1], F[1], atol=1e-15)


def test_rest_state_is_invariant():
    c = R.Case((10, 8), tau=0.9)
    f0 = R.initial_populations(c)
    assert np.allclose(R.run(c, 50, f0), f0, atol=1e-15)


def test_constant_force_field_equals_uniform_force():
    F = (2e-5, -1e-5)
    ff = np.stack([np.full((6, 9), F[0]), np.full((6, 9), F[1])])
    for coll in ("bgk", "trt", "mrt"):
        a = R.run(R.Case((9, 6), tau=0.7, periodic=(True, False), force=F, collision=coll), 30)
        b = R.run(R.Case((9, 6), tau=0.7, periodic=(True, False), force_field=ff,
                         collision=coll), 30)
        assert np.array_equal(a, b)


# ----------------------------------------------------------------------------- streaming


def _brute_stream(f, case):
    """Per node and direction: source x - c_i; any wall crossing -> bounce-back of the
    own opp(i) value, plus the Ladd term if the crossing is through the lid."""
    lat, d = case.lat, case.dim
    out = np.empty_like(f)
    for i in range(lat.q):
        for idx in itertools.product(*[range(s) for s in case.array_shape]):
            x = [idx[d - 1 - a] for a in range(d)]
            src, wall, lid = [], False, False
            for a in range(d):
                s = x[a] - lat.c[i, a]
                if s < 0 or s >= case.shape[a]:
                    if case.periodic[a]:
                        s %= case.shape[a]
                    else:
                        wall = True
                        lid = lid or (a == d - 1 and s >= case.shape[a])
                src.append(s)
            if wall:
                v = f[(lat.opp[i],) + idx]
                if lid and case.u_lid != 0.0:
                    v = v + 6.0 * lat.w[i] * lat.c[i, 0] * case.u_lid
            else:
                v = f[(i,) + tuple(src[d - 1 - k] for k in range(d))]
            out[(i,) + idx] = v
    return out


@pytest.mark.parametrize("name", list(STREAM_CASES))
def test_stream_equals_brute_force_rule(units, name):
    case = R.Case(**STREAM_CASES[name])
    f = units[f"{name}/in"]
    assert np.array_equal(R.stream(f, case), _brute_stream(f, case))


def test_periodic_stream_is_a_permutation():
    c = R.Case((5, 4, 3), periodic=(True, True, True))
    f = _rand_f(R.D3Q19, (3, 4, 5), 4)
    g = R.stream(f, c)
    for i in range(19):
        assert np.array_equal(np.sort(g[i].ravel()), np.sort(f[i].ravel()))


def test_closed_box_conserves_mass():
    c = R.Case((24, 20), tau=3.0 * 0.1 * 24 / 50 + 0.5, u_lid=0.1)
    f0 = R.initial_populations(c)
    f = R.run(c, 300, f0)
    assert abs(f.sum() - f0.sum()) < 1e-10 * f0.sum()


def test_channel_is_symmetric():
    c = R.Case((6, 17), tau=0.9, periodic=(True, False), force=(1e-5, 0.0))
    f = R.run(c, 400)
    _, u = R.macroscopic(f, c)
    assert np.allclose(u[0], u[0][::-1], atol=1e-14)
    assert np.abs(u[1]).max() < 1e-14


# ----------------------------------------------------------------------------- AA + storage

AA_CASES = {
    "cavity2d-trt": dict(shape=(9, 7), tau=0.6, u_lid=0.1, collision="trt"),
    "channel-mrt-force": dict(shape=(8, 6), tau=0.7, periodic=(True, False),
                              force=(1e-4, -2e-5), collision="mrt"),
    "cavity3d-bgk": dict(shape=(5, 4, 6), tau=0.7, u_lid=0.08),
    "mixed3d-trt-force": dict(shape=(5, 6, 4), tau=0.8, periodic=(True, False, True),
                              force=(1e-4, 0.0, 2e-5), collision="trt"),
}


@pytest.mark.parametrize("name", list(AA_CASES))
def test_aa_pairs_equal_two_buffer_steps(name):
    c = R.Case(**AA_CASES[name])
    f0 = R.initial_populations(c) * (1 + 0.01 * np.random.default_rng(5).standard_normal(
        (c.lat.q,) + c.array_shape))
    A = f0[c.lat.opp]  # swapped layout: slot opp(i) holds population i
    f = f0
    for _ in range(4):
        A = R.aa_even(R.aa_odd(A, c), c)
        f = R.step(R.step(f, c), c)
        assert np.array_equal(A[c.lat.opp], f)


def test_aa_odd_leaves_natural_streamed_layout():
    """After an odd step, slot i at x holds the population i that streamed into x."""
    c = R.Case((6, 5), tau=0.7, periodic=(True, True))
    f0 = R.initial_populations(c) * (1 + 0.01 * np.random.default_rng(6).standard_normal(
        (9, 5, 6)))
    A = R.aa_odd(f0[c.lat.opp], c)
    assert np.array_equal(A, R.stream(R.collide(R.stream(f0, c), c), c))


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(7)
    x = (rng.standard_normal(100000) * 10.0 ** rng.uniform(-8, 2, 100000)).astype(np.float32)
    x[:4] = [1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -(1.0 + 2 ** -8), 0.0]  # ties
    ours = R.bf16_rne_bits(x)
    theirs = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(ours, theirs)


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("aa", [False, True])
def test_encode_decode_roundtrip(storage, aa):
    lat = R.D2Q9
    f = _rand_f(lat, (3, 4), 8)
    g = R.decode(R.encode(f, storage, aa, lat), storage, aa, lat)
    tol = {"f32": 1e-7, "f16": 2e-4, "bf16": 2e-3}[storage]
    assert np.abs(g - f).max() < tol * np.abs(f).max()
    # shifted storage keeps f - w, so the rest state is stored exactly as zero
    rest = R.equilibrium(lat, np.ones((3, 4)), np.zeros((2, 3, 4)))
    if storage != "f32":
        assert not np.any(R.encode(rest, storage, aa, lat).astype(np.float64))


@pytest.mark.parametrize("dim,periodic,u_lid,force", [
    (2, (True, True), 0.0, None), (2, (False, False), 0.05, None),
    (2, (True, False), 0.0, (1e-5, 0.0)), (3, (False, True, False), 0.05, None)])
def test_aa_launches_ignore_a_ghost_offset_at_tau_one(dim, periodic, u_lid, force):
    """Premise of the GPU write-every-slot test of the AA kernels: with BGK at tau = 1 an AA
    launch maps X and X + eps g (g the constant ghost field of tests/test_hip_poison.py) to
    the same array, while leaving a slot unwritten would keep its offset eps g_i."""
    from tests.test_hip_poison import _ghost
    from pylbm_amd.lattice import LATTICES

    shape = (9, 7) if dim == 2 else (6, 5, 4)
    case = R.Case(shape=shape, tau=1.0, periodic=periodic, u_lid=u_lid, force=force)
    lat = case.lat
    g = _ghost(LATTICES["D2Q9" if dim == 2 else "D3Q19"])
    rng = np.random.default_rng(1)
    X = R.equilibrium(lat, 1 + 0.01 * rng.standard_normal(case.array_shape),
                      [0.03 * rng.standard_normal(case.array_shape) for _ in range(dim)])
    X = X + 0.005 * rng.standard_normal(X.shape)
    Xg = X + 0.025 * g.reshape((-1,) + (1,) * dim)
    for launch in (R.aa_odd, R.aa_even):
        a, b = launch(X, case), launch(Xg, case)
        assert np.abs(a - b).max() < 1e-14
        # a launch that left a slot unwritten would return the two inputs there, which
        # differ by eps |g_i| >= 0.025 in every slot
        assert np.abs(Xg - X).min() >= 0.025 - 1e-12
