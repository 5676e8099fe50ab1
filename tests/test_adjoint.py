"""Gradients of the checkpointed rollout (HIP adjoint) vs float64 references.

There is no reference gradient path (the reference's torch rungs run under no_grad), so
the adjoint is pinned by: (a) float64 torch autograd through oracle/torch_ref.py, a
restatement of the step that tests below tie to the NumPy oracle; (b) central finite
differences of the float64 C oracle; (c) a per-step VJP comparison on random states.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import lbm_ref as R
from oracle import lbm_ref_c as RC
from oracle.torch_ref import TorchLBM
from pylbm_amd.adjoint import tau_rate_derivatives
from pylbm_amd.fields import divergence_free_field, smooth_random_field
from pylbm_amd.problem import Problem

# ----------------------------------------------------------------------------- CPU


@pytest.mark.parametrize("kw", [
    dict(shape=(9, 7), tau=0.8, periodic=(True, False), collision="bgk"),
    dict(shape=(9, 7), tau=0.7, u_lid=0.05, collision="trt"),
    dict(shape=(5, 4, 6), tau=0.75, periodic=(True, False, True), collision="trt"),
])
def test_torch_restatement_equals_oracle(kw):
    rng = np.random.default_rng(0)
    case = R.Case(**kw)
    ff = 1e-4 * rng.standard_normal((case.dim,) + case.array_shape)
    case_f = R.Case(**kw, force_field=ff)
    f0 = R.initial_populations(case) * (1 + 0.02 * rng.standard_normal(
        (case.lat.q,) + case.array_shape))
    tl = TorchLBM(case_f)
    got = tl.rollout(torch.tensor(f0), torch.tensor(ff), torch.tensor(case.tau, dtype=torch.float64),
                     5).numpy()
    assert np.abs(got - R.run(case_f, 5, f0)).max() < 1e-15


def test_torch_restatement_gradient_matches_finite_differences():
    case = R.Case((8, 6), tau=0.8, periodic=(True, False), collision="trt")
    rng = np.random.default_rng(1)
    ff = torch.tensor(1e-5 + 1e-6 * rng.standard_normal((2, 6, 8)), requires_grad=True)
    tau = torch.tensor(0.8, dtype=torch.float64, requires_grad=True)
    tl = TorchLBM(case)
    ut = torch.tensor(0.001 * rng.standard_normal((2, 6, 8)))

    def loss(F, t):
        f = tl.rollout(tl.equilibrium(torch.ones(6, 8, dtype=torch.float64),
                                      torch.zeros(2, 6, 8, dtype=torch.float64)), F, t, 30)
        return 0.5 * ((tl.velocity(f, F) - ut) ** 2).sum()

    L = loss(ff, tau)
    gF, gt = torch.autograd.grad(L, (ff, tau))
    h = 1e-4 * 0.8
    fd = (loss(ff.detach(), tau.detach() + h) - loss(ff.detach(), tau.detach() - h)) / (2 * h)
    assert abs(float(gt) - float(fd)) < 1e-6 * abs(float(fd))
    V = torch.tensor(rng.standard_normal((2, 6, 8)))
    hF = 1e-4 * 1e-5
    fdF = (loss(ff.detach() + hF * V, tau.detach()) - loss(ff.detach() - hF * V, tau.detach())) \
        / (2 * hF)
    assert abs(float((gF * V).sum()) - float(fdF)) < 1e-6 * abs(float(fdF))


@pytest.mark.parametrize("lattice,shape,periodic", [
    ("D2Q9", (9, 6), (False, False)), ("D2Q9", (8, 5), (True, False)),
    ("D3Q19", (7, 5, 4), (False, True, False)), ("D3Q19", (6, 5, 4), (True, True, True))])
def test_inplace_adjoint_index_rule(lattice, shape, periodic):
    """The layout rule of the in-place (AA) adjoint, on the CPU: an even step (own slots in,
    mu_k to the own slot opp(k)) and an odd step (lam_i(x) from (opp(i), x + c_i), or the
    own slot i when x + c_i is behind a wall; mu pushed as the two-buffer step pushes)
    equal two two-buffer push steps, with any per-node map standing in for the transposed
    collision, and each node of the odd step writes exactly the slots it reads."""
    from pylbm_amd.lattice import LATTICES

    lat = LATTICES[lattice]
    c, opp, q, d = np.array(lat.c), np.array(lat.opp), lat.q, len(shape)
    rng = np.random.default_rng(1)
    M = rng.standard_normal((q, q))
    nodes = list(np.ndindex(*shape))

    def inside(x):
        y = list(x)
        for a in range(d):
            if not 0 <= y[a] < shape[a]:
                if not periodic[a]:
                    return None
                y[a] %= shape[a]
        return tuple(y)

    def at(k, x):
        return (k,) + tuple(reversed(x))

    def mu(lam, x):
        return M @ lam + 0.01 * sum(x)

    def push(out, x, m, touched=None):
        for k in range(q):
            src = inside(tuple(np.array(x) - c[k, :d]))
            slot = at(opp[k], x) if src is None else at(k, src)
            out[slot] = m[k]
            if touched is not None:
                touched.append(slot)

    def ab(lam):
        out = np.full_like(lam, np.nan)
        for x in nodes:
            push(out, x, mu(lam[(slice(None),) + tuple(reversed(x))], x))
        return out

    lam0 = rng.standard_normal((q,) + tuple(reversed(shape)))
    A = lam0.copy()
    for x in nodes:  # even
        m = mu(lam0[(slice(None),) + tuple(reversed(x))], x)
        for k in range(q):
            A[at(opp[k], x)] = m[k]
    B = A.copy()
    for x in nodes:  # odd
        reads, writes, lam = [], [], np.empty(q)
        for i in range(q):
            nb = inside(tuple(np.array(x) + c[i, :d]))
            reads.append(at(i, x) if nb is None else at(opp[i], nb))
            lam[i] = A[reads[-1]]
        push(B, x, mu(lam, x), writes)
        assert sorted(reads) == sorted(writes)
    np.testing.assert_array_equal(B, ab(ab(lam0)))


def test_rollout_config_checks_adjoint_storage():
    from pylbm_amd.adjoint import RolloutConfig

    p = Problem((8, 8), tau=0.8)
    assert RolloutConfig(p, 4, lam_storage="f16").lam_storage == "f16"
    with pytest.raises(ValueError, match="lam_storage"):
        RolloutConfig(p, 4, lam_storage="bf16")
    with pytest.raises(ValueError, match="BGK and TRT"):
        RolloutConfig(Problem((8, 8), tau=0.8, collision="mrt"), 4)
    assert RolloutConfig(p, 4).records is None  # records unless tau is differentiated


def test_lam16_scale_is_a_power_of_two_near_the_target():
    from pylbm_amd.adjoint import lam16_scale

    for g in (1e-9, 3e-4, 0.03, 1.0, 7.5e3, 1e12):
        s = lam16_scale(g)
        assert s == 2.0 ** round(np.log2(s))
        assert 2.0 ** 7 <= s * g <= 2.0 ** 9
    assert lam16_scale(0.0) == 1.0 and lam16_scale(float("nan")) == 1.0
    assert lam16_scale(float("inf")) == 1.0


@pytest.mark.parametrize("coll", ["bgk", "trt"])
def test_tau_rate_derivatives(coll):
    p = Problem((4, 4), tau=0.8, collision=coll)
    a, b = tau_rate_derivatives(p, 0.8)
    h = 1e-6

    def rates(t):
        return Problem((4, 4), tau=t, collision=coll).runtime_rates()

    fd = [(rp - rm) / (2 * h) for rp, rm in zip(rates(0.8 + h), rates(0.8 - h))]
    assert np.isclose(a, fd[0], rtol=1e-7)
    assert np.isclose(b, fd[1] if coll == "trt" else 0.0, rtol=1e-7, atol=1e-12)


# ----------------------------------------------------------------------------- GPU


def _dev(a):
    return torch.as_tensor(a, device="cuda")


VJP_CASES = {
    "d2q9-bgk-channel": dict(shape=(24, 16), tau=0.8, periodic=(True, False), collision="bgk"),
    "d2q9-trt-cavity": dict(shape=(20, 12), tau=0.7, u_lid=0.05, collision="trt"),
    "d3q19-bgk-periodic": dict(shape=(12, 8, 6), tau=0.7, periodic=(True, True, True)),
    "d3q19-trt-walls": dict(shape=(10, 8, 6), tau=0.65, periodic=(True, False, False),
                            collision="trt", u_lid=0.03),
    # rows of 600 nodes: two nodes per lane, a partial second segment, the x wrap on both
    "d2q9-trt-wide-rows": dict(shape=(600, 5), tau=0.75, periodic=(True, False),
                               collision="trt", u_lid=0.02),
    "d2q9-bgk-wide-walls": dict(shape=(530, 4), tau=0.8, collision="bgk", u_lid=0.04),
}


@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["f32", "f32s"])
@pytest.mark.parametrize("name", list(VJP_CASES))
def test_adjoint_step_equals_float64_vjp(name, storage):
    """One adjoint step (lam, dF, dL/dtau) == torch float64 VJP of one restated step."""
    from pylbm_amd import _lib
    from pylbm_amd.adjoint import RolloutConfig, _Engine

    kw = VJP_CASES[name]
    rng = np.random.default_rng(5)
    case = R.Case(**kw)
    q, d, shp = case.lat.q, case.dim, case.array_shape
    ff = (1e-4 * rng.standard_normal((d,) + shp)).astype(np.float32).astype(np.float64)
    f = (R.initial_populations(case) * (1 + 0.02 * rng.standard_normal((q,) + shp)))
    f = f.astype(np.float32).astype(np.float64)
    lam = rng.standard_normal((q,) + shp).astype(np.float32).astype(np.float64)
    # float64 reference
    tl = TorchLBM(R.Case(**kw, force_field=ff))
    ft = torch.tensor(f, requires_grad=True)
    Ft = torch.tensor(ff, requires_grad=True)
    tt = torch.tensor(case.tau, dtype=torch.float64, requires_grad=True)
    out = tl.step(ft, Ft, tt)
    gf, gF, gt = torch.autograd.grad((out * torch.tensor(lam)).sum(), (ft, Ft, tt))
    # HIP
    p = Problem(**kw, force_field=ff)
    eng = _Engine(RolloutConfig(p, 1, storage=storage),
                  _dev(ff.astype(np.float32)).reshape(d, -1), case.tau)
    stored = R.encode(f, storage, False, case.lat)  # f32: f; f32s: f - w, both in fp32
    fs = _dev(stored).reshape(-1)
    lin = _dev(lam.astype(np.float32)).reshape(-1)
    lout = torch.full_like(lin, float("nan"))
    dF = torch.zeros(d * p.n_nodes, dtype=torch.float32, device="cuda")
    part = torch.zeros(int(_lib.load().pylbm_hip_adjoint_workspace(eng.ref)), dtype=torch.float64,
                       device="cuda")
    _lib.call("pylbm_hip_adjoint_step", eng.ref, fs.data_ptr(), lin.data_ptr(),
              lout.data_ptr(), dF.data_ptr(), part.data_ptr(), None)
    torch.cuda.synchronize()
    got_f = lout.double().cpu().numpy().reshape(gf.shape)
    assert np.isfinite(got_f).all(), "every adjoint slot must be written exactly once"
    rel = lambda a, b: np.abs(a - b).max() / np.abs(b).max()  # noqa: E731
    e_f = rel(got_f, gf.numpy())
    e_F = rel(dF.double().cpu().numpy().reshape(gF.shape), gF.numpy())
    a, b = tau_rate_derivatives(p, case.tau)
    sums = part.view(-1, 2).sum(0).cpu().numpy()
    e_t = abs(sums[0] * a + sums[1] * b - float(gt)) / abs(float(gt))
    print(f"{name} {storage}: rel err lam {e_f:.2e}  dF {e_F:.2e}  dtau {e_t:.2e}")
    # measured: lam <= 2.0e-7, dF <= 2.0e-7, dtau <= 8.9e-6
    assert e_f < 2e-6 and e_F < 2e-6 and e_t < 2e-5


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(VJP_CASES))
def test_adjoint_step_fp16_states(name):
    """fp16 adjoint states: the step computes in fp32 exactly as with fp32 states and only
    rounds lam_out to fp16, so on fp16-representable lam_in it returns the fp32-state
    kernel's lam_out rounded to nearest fp16 (within one fp16 ulp where the fp32 results sit
    next to a rounding tie), the same dF and the same dL/ds partial sums."""
    from pylbm_amd import _lib
    from pylbm_amd.adjoint import RolloutConfig, _Engine

    kw = VJP_CASES[name]
    rng = np.random.default_rng(6)
    case = R.Case(**kw)
    q, d, shp = case.lat.q, case.dim, case.array_shape
    ff = (1e-4 * rng.standard_normal((d,) + shp)).astype(np.float32)
    f = (R.initial_populations(case) * (1 + 0.02 * rng.standard_normal((q,) + shp)))
    lam = (0.05 * rng.standard_normal((q,) + shp)).astype(np.float16)
    p = Problem(**kw, force_field=ff.astype(np.float64))
    fs = _dev(R.encode(f.astype(np.float32).astype(np.float64), "f32s", False,
                       case.lat)).reshape(-1)
    out = {}
    for ls in ("f32", "f16"):
        eng = _Engine(RolloutConfig(p, 1, storage="f32s", lam_storage=ls),
                      _dev(ff).reshape(d, -1), case.tau)
        dt = torch.float16 if ls == "f16" else torch.float32
        lin = _dev(lam).to(dt).reshape(-1)
        lout = torch.full_like(lin, float("nan"))
        dF = torch.zeros(d * p.n_nodes, dtype=torch.float32, device="cuda")
        part = torch.zeros(int(_lib.load().pylbm_hip_adjoint_workspace(eng.ref)),
                           dtype=torch.float64, device="cuda")
        _lib.call("pylbm_hip_adjoint_step", eng.ref, fs.data_ptr(), lin.data_ptr(),
                  lout.data_ptr(), dF.data_ptr(), part.data_ptr(), None)
        torch.cuda.synchronize()
        out[ls] = (lout.cpu(), dF.cpu().numpy(), part.cpu().numpy())
    l16, l32 = out["f16"][0], out["f32"][0]
    assert torch.isfinite(l16).all(), "every adjoint slot must be written exactly once"
    r16 = l32.to(torch.float16)
    exact = (l16 == r16).double().mean().item()
    # one fp16 ulp at the rounded value (2^-24 in the subnormal range)
    a16 = r16.float().abs()
    ulp16 = torch.where(a16 >= 2.0 ** -14,
                        2.0 ** (torch.floor(torch.log2(a16.clamp_min(2.0 ** -14))) - 10),
                        torch.full_like(a16, 2.0 ** -24))
    worst = ((l16.float() - r16.float()).abs() / ulp16).max().item()
    print(f"{name}: fp16 states == rounded fp32 states at {100 * exact:.3f}% of slots, "
          f"worst {worst:.1f} ulp")
    assert worst <= 1.0
    assert np.allclose(out["f16"][1], out["f32"][1], rtol=1e-6, atol=1e-12)
    assert np.allclose(out["f16"][2], out["f32"][2], rtol=1e-6, atol=1e-12)


def _channel_setup(nx, ny, seed=4):
    rng = np.random.default_rng(seed)
    ff = np.zeros((2, ny, nx))
    ff[0] = 1e-6
    ff += 1e-7 * rng.standard_normal((2, ny, nx))
    ut = smooth_random_field((nx, ny), 4, 0.01, seed + 1)
    return ff.astype(np.float32).astype(np.float64), ut


@pytest.mark.gpu
def test_channel_gradients_vs_float64_autograd_and_fd():
    """D2Q9 BGK channel 128x64, tau = 0.8, T = 200, per-cell force: dL/dF and dL/dtau."""
    from pylbm_amd.adjoint import RolloutConfig, rollout, velocity_loss

    nx, ny, T = 128, 64, 200
    ff, ut = _channel_setup(nx, ny)
    kw = dict(shape=(nx, ny), tau=0.8, periodic=(True, False), collision="bgk")
    p = Problem(**kw, force_field=ff)
    cfg = RolloutConfig(p, T, checkpoint=50)
    F = _dev(ff).float().requires_grad_(True)
    tau = torch.tensor(0.8, dtype=torch.float64, device="cuda", requires_grad=True)
    L = velocity_loss(rollout(cfg, F, tau), _dev(ut))
    L.backward()
    gF_hip = F.grad.double().cpu().numpy()
    gt_hip = float(tau.grad)
    # float64 autograd of the restated step
    tl = TorchLBM(R.Case(**kw, force_field=ff))
    Ft = torch.tensor(ff, requires_grad=True)
    tt = torch.tensor(0.8, dtype=torch.float64, requires_grad=True)
    f0 = tl.equilibrium(torch.ones(ny, nx, dtype=torch.float64),
                        torch.zeros(2, ny, nx, dtype=torch.float64))
    fT = tl.rollout(f0, Ft, tt, T, segment=50)
    Lt = 0.5 * ((tl.velocity(fT, Ft) - torch.tensor(ut)) ** 2).sum()
    gF, gt = torch.autograd.grad(Lt, (Ft, tt))
    eF = np.abs(gF_hip - gF.numpy()).max() / np.abs(gF.numpy()).max()
    et = abs(gt_hip - float(gt)) / abs(float(gt))
    # central finite differences (h = 1e-4 relative) of the float64 C oracle
    def loss64(tau_, ffield):
        c = R.Case(**{**kw, "tau": tau_}, force_field=ffield)
        f = RC.run(c, T)
        _, u = R.macroscopic(f, c)
        return 0.5 * float(((u - ut) ** 2).sum())

    h = 1e-4 * 0.8
    fd_t = (loss64(0.8 + h, ff) - loss64(0.8 - h, ff)) / (2 * h)
    V = np.random.default_rng(9).standard_normal(ff.shape)
    hF = 1e-4 * np.abs(ff).max()
    fd_F = (loss64(0.8, ff + hF * V) - loss64(0.8, ff - hF * V)) / (2 * hF)
    dd = float((gF_hip * V).sum())
    print(f"channel: L={float(L.detach()):.6e} (f64 {float(Lt.detach()):.6e}); "
          f"dF rel err vs autograd {eF:.2e}; "
          f"dtau {gt_hip:.6e} vs autograd {float(gt):.6e} ({et:.2e}) vs FD {fd_t:.6e}; "
          f"<dF,V> {dd:.6e} vs FD {fd_F:.6e}")
    # measured: dF 3.0e-6 and dtau 2.0e-5 vs autograd; dtau 2.0e-5, <dF,V> 1.7e-6 vs FD
    assert abs(float(L.detach()) - float(Lt.detach())) < 1e-5 * float(Lt.detach())
    assert eF < 3e-5 and et < 1e-4
    assert abs(gt_hip - fd_t) < 1e-4 * abs(fd_t)
    assert abs(dd - fd_F) < 2e-5 * abs(fd_F)


@pytest.mark.gpu
def test_channel_gradients_with_fp16_adjoint_states():
    """fp16 adjoint states on the force-field channel (128x64, T = 200): the rollout scales
    the terminal gradient by a power of two for fp16's range and divides dL/dF and dL/dtau
    by it again; both stay within the fp16 rounding of the fp32-state gradients."""
    from pylbm_amd.adjoint import RolloutConfig, lam16_scale, rollout, velocity_loss

    nx, ny, T = 128, 64, 200
    ff, ut = _channel_setup(nx, ny)
    kw = dict(shape=(nx, ny), tau=0.8, periodic=(True, False), collision="bgk")
    p = Problem(**kw, force_field=ff)
    got = {}
    for ls in ("f32", "f16"):
        cfg = RolloutConfig(p, T, checkpoint=50, lam_storage=ls)
        F = _dev(ff).float().requires_grad_(True)
        tau = torch.tensor(0.8, dtype=torch.float64, device="cuda", requires_grad=True)
        velocity_loss(rollout(cfg, F, tau), _dev(ut)).backward()
        got[ls] = (F.grad.double().cpu().numpy(), float(tau.grad))
    eF = np.abs(got["f16"][0] - got["f32"][0]).max() / np.abs(got["f32"][0]).max()
    et = abs(got["f16"][1] - got["f32"][1]) / abs(got["f32"][1])
    print(f"channel fp16 adjoint states: dF rel err {eF:.2e}, dtau rel err {et:.2e} "
          f"(scale {lam16_scale(1e-2):g} for |dL/du| = 1e-2)")
    assert eF < 1.2e-3 and et < 2e-3  # measured 1.1e-4 and 7.7e-4


@pytest.mark.gpu
@pytest.mark.parametrize("lam_storage", ["f32", "f16"])
def test_overlapped_recompute_gives_identical_gradients(lam_storage):
    """The backward pass recomputes segment k - 1 on a second stream during the adjoint of
    segment k; the adjoint steps stay in order on one stream, so the gradients equal the
    one-stream backward pass bit for bit (T = 250, C = 40: 7 segments, a short last one)."""
    from pylbm_amd.adjoint import RolloutConfig, rollout, velocity_loss

    nx, ny, T = 96, 40, 250
    ff, ut = _channel_setup(nx, ny, seed=7)
    kw = dict(shape=(nx, ny), tau=0.75, periodic=(True, False), collision="trt")
    p = Problem(**kw, force_field=ff)
    grads = []
    for overlap in (True, False, True):
        cfg = RolloutConfig(p, T, checkpoint=40, lam_storage=lam_storage, overlap=overlap)
        F = _dev(ff).float().requires_grad_(True)
        tau = torch.tensor(0.75, dtype=torch.float64, device="cuda", requires_grad=True)
        velocity_loss(rollout(cfg, F, tau), _dev(ut)).backward()
        grads.append((F.grad.cpu().numpy(), float(tau.grad)))
    for g in grads[1:]:
        assert np.array_equal(g[0], grads[0][0]) and g[1] == grads[0][1]


@pytest.mark.gpu
@pytest.mark.parametrize("storage,lam_storage,bound", [
    # bounds about 10x the measured errors: f32s 4.9e-6, f32 4.7e-6, f16 forward 3.1e-5
    ("f32s", "f32", 5e-5), ("f32", "f32", 5e-5), ("f16", "f32", 3e-4),
    # fp16 adjoint states: measured 2.9e-4 (f32s) and 2.5e-4 (f16); the float64 adjoint
    # with its states rounded to fp16 after every step measured 3.1e-4 (the rounding alone)
    ("f32s", "f16", 3e-3), ("f16", "f16", 3e-3)])
def test_d3q19_initial_velocity_gradient(storage, lam_storage, bound):
    """D3Q19 BGK periodic 32^3, tau = 0.7: dL/du0 vs float64 autograd (T = 100)."""
    from pylbm_amd.adjoint import RolloutConfig, rollout, velocity_loss

    n, T = 32, 100
    u0 = divergence_free_field(n, 8, 0.02, 1)
    uo = divergence_free_field(n, 8, 0.02, 2)
    kw = dict(shape=(n, n, n), tau=0.7, periodic=(True, True, True))
    cfg = RolloutConfig(Problem(**kw), T, checkpoint=25, storage=storage,
                        lam_storage=lam_storage)
    U0 = _dev(u0).requires_grad_(True)
    L = velocity_loss(rollout(cfg, u0=U0), _dev(uo))
    L.backward()
    g_hip = U0.grad.cpu().numpy()
    tl = TorchLBM(R.Case(**kw))
    Ut = torch.tensor(u0, requires_grad=True)
    zero = torch.zeros(3, n, n, n, dtype=torch.float64)
    fT = tl.rollout(tl.equilibrium(torch.ones(n, n, n, dtype=torch.float64), Ut), zero,
                    torch.tensor(0.7, dtype=torch.float64), T, segment=25)
    Lt = 0.5 * ((tl.velocity(fT, zero) - torch.tensor(uo)) ** 2).sum()
    (g,) = torch.autograd.grad(Lt, (Ut,))
    err = np.abs(g_hip - g.numpy()).max() / np.abs(g.numpy()).max()
    print(f"D3Q19 32^3 T={T} {storage} lam {lam_storage}: L {float(L):.6e} vs f64 "
          f"{float(Lt):.6e}; "
          f"du0 rel err {err:.2e} (bound {bound:.0e})")
    assert err < bound


@pytest.mark.gpu
def test_channel_gradients_at_full_size_vs_central_differences():
    """The differentiable channel at its stated size: D2Q9 BGK 1024 x 256, tau = 0.8, per-cell
    force (1e-6 + 1e-7 N(0,1)), T = 2000, checkpoints every 100, f32s storage.

    * The gradients are finite; the overlapped and one-stream backward passes agree bit for
      bit.
    * dL/dtau against central differences of the same HIP forward, i.e. the derivative of
      the model the adjoint differentiates: steps of 1% and 3% of tau, Richardson-
      extrapolated (a smaller step drowns in the fp32 forward's rounding).  Bound 1e-4;
      measured 4.8e-6.
    * Against the float64 oracle's central differences (tests/golden/channel_fd64.json,
      two step sizes each, agreeing to 1e-6): the loss agrees to 6.4e-6 and <dL/dF, V> to
      3.7e-5 (bound 2e-4).  dL/dtau differs by 3.3e-4 (bound 1e-3): that is the
      difference between the fp32 and the float64 trajectories' tau sensitivity over 2000
      steps (fp32 instead of f32s storage alone moves it by 1.5e-4), not an adjoint error
      (the check above).  <dL/dF, V> is not compared with differences of the HIP forward:
      its loss carries rounding of ~1e-7 of L, more than the change of any force step
      small enough for a central difference (a 10% step left 6.5e-4).
    """
    import json
    from pathlib import Path

    from pylbm_amd.adjoint import RolloutConfig, rollout, velocity_loss

    ref = json.loads((Path(__file__).parent / "golden" / "channel_fd64.json").read_text())
    nx, ny, T, C = 1024, 256, 2000, 100
    ff, ut = _channel_setup(nx, ny, seed=ref["seed"])
    kw = dict(shape=(nx, ny), tau=0.8, periodic=(True, False), collision="bgk")
    p = Problem(**kw, force_field=ff)
    target = _dev(ut)
    grads = {}
    for overlap in (True, False):
        cfg = RolloutConfig(p, T, checkpoint=C, overlap=overlap)
        F = _dev(ff).float().requires_grad_(True)
        tau = torch.tensor(0.8, dtype=torch.float64, device="cuda", requires_grad=True)
        L = velocity_loss(rollout(cfg, F, tau), target)
        L.backward()
        grads[overlap] = (F.grad.clone(), float(tau.grad), float(L.detach()))
    gF, gt, L0 = grads[True]
    assert bool(torch.isfinite(gF).all()) and np.isfinite(gt)
    assert torch.equal(gF, grads[False][0]) and gt == grads[False][1]
    cfg = RolloutConfig(p, T, checkpoint=C)

    def loss(tau_: float, field: torch.Tensor) -> float:
        with torch.no_grad():
            u = rollout(cfg, field.float(), torch.tensor(tau_, dtype=torch.float64,
                                                         device="cuda"))
            return float(velocity_loss(u, target))

    F0 = _dev(ff)
    fd = [(loss(0.8 + h, F0) - loss(0.8 - h, F0)) / (2 * h) for h in (0.008, 0.024)]
    fd_t = (9 * fd[0] - fd[1]) / 8  # central differences err by c h^2: eliminate it
    V = torch.as_tensor(np.random.default_rng(ref["v_seed"]).standard_normal(ff.shape),
                        device="cuda")
    dd = float((gF.double() * V).sum())
    e_t = abs(gt - fd_t) / abs(fd_t)
    t64, F64 = ref["dL_dtau"]["0.0001"], ref["dL_dF_dot_V"]["0.01"]
    e_L64 = abs(L0 - ref["loss"]) / ref["loss"]
    e_t64, e_F64 = abs(gt - t64) / abs(t64), abs(dd - F64) / abs(F64)
    print(f"channel 1024x256 T={T}: L {L0:.8e} (float64 {ref['loss']:.8e}, {e_L64:.1e})\n"
          f"  dL/dtau  adjoint {gt:.8e}; HIP FD {fd_t:.8e} ({e_t:.1e}); "
          f"float64 FD {t64:.8e} ({e_t64:.1e})\n"
          f"  <dL/dF,V> adjoint {dd:.8e}; float64 FD {F64:.8e} ({e_F64:.1e})")
    assert e_t < 1e-4
    assert e_L64 < 5e-5 and e_F64 < 2e-4 and e_t64 < 1e-3
