"""Moment records (GPU): the adjoint of rollouts that do not differentiate tau.

The forward step writes, per node, the density and velocity its collision used
(pylbm_hip_run_record), and the adjoint takes them from there instead of pulling the stored
state again (pylbm_hip_adjoint_step_record / pylbm_hip_adjoint_run_record).  Pinned by:
the records against float64 moments of the oracle's pulled populations; one record-mode
adjoint step against the torch float64 VJP of the restated step; the record-mode step
against the state-mode step; the in-place record run against single two-buffer record
steps, bit for bit; and rollout gradients with records against those with stored states
and against float64 autograd.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from oracle import lbm_ref as R
from oracle.torch_ref import TorchLBM
from pylbm_amd.fields import divergence_free_field, smooth_random_field
from pylbm_amd.problem import Problem

pytestmark = pytest.mark.gpu

CASES = {
    "d2q9-bgk-channel-field": dict(shape=(24, 16), tau=0.8, periodic=(True, False),
                                   collision="bgk"),
    "d2q9-trt-cavity": dict(shape=(20, 12), tau=0.7, u_lid=0.05, collision="trt"),
    "d3q19-bgk-periodic": dict(shape=(12, 8, 6), tau=0.7, periodic=(True, True, True)),
    "d3q19-trt-walls": dict(shape=(10, 8, 6), tau=0.65, periodic=(True, False, False),
                            collision="trt", u_lid=0.03),
    # rows of 600 nodes: several workgroup segments, a partial last one, the x wrap
    "d2q9-bgk-wide-rows": dict(shape=(600, 5), tau=0.75, periodic=(True, False),
                               collision="bgk", u_lid=0.02),
}


def _dev(a):
    return torch.as_tensor(a, device="cuda")


def _setup(name, storage, seed):
    """Problem with a per-cell force, a random stored state f^t, its float64 value, engine."""
    from pylbm_amd.adjoint import RolloutConfig, _Engine

    kw = CASES[name]
    rng = np.random.default_rng(seed)
    case = R.Case(**kw)
    q, d, shp = case.lat.q, case.dim, case.array_shape
    ff = (1e-4 * rng.standard_normal((d,) + shp)).astype(np.float32).astype(np.float64)
    f = R.initial_populations(case) * (1 + 0.02 * rng.standard_normal((q,) + shp))
    stored = R.encode(f, storage, False, case.lat)
    f64 = R.decode(stored, storage, False, case.lat)  # the value the kernels see
    p = Problem(**kw, force_field=ff)
    eng = _Engine(RolloutConfig(p, 1, storage=storage), _dev(ff.astype(np.float32)).reshape(d, -1),
                  case.tau)
    return kw, case, ff, f64, _dev(stored).reshape(-1), p, eng, rng


def _record(eng, fs):
    from pylbm_amd import _lib
    from pylbm_amd.adjoint import _ptr_array

    dst = eng.state()
    rec = torch.full((4 * eng.N,), float("nan"), dtype=torch.float32, device="cuda")
    _lib.call("pylbm_hip_run_record", eng.ref, fs.data_ptr(), _ptr_array([dst]),
              _ptr_array([rec]), 1, None)
    return dst, rec


@pytest.mark.parametrize("storage", ["f32", "f16"])
@pytest.mark.parametrize("name", list(CASES))
def test_records_are_the_moments_of_the_pulled_populations(name, storage):
    """rho and u = (sum c g + F/2) / rho of the pulled populations g (walls, lid, force),
    against float64 moments of the oracle's stream of the same stored state; the state the
    step writes equals pylbm_hip_run_store's bit for bit."""
    from pylbm_amd import _lib
    from pylbm_amd.adjoint import _ptr_array

    kw, case, ff, f64, fs, p, eng, _ = _setup(name, storage, 11)
    dst, rec = _record(eng, fs)
    plain = eng.state()
    _lib.call("pylbm_hip_run_store", eng.ref, fs.data_ptr(), _ptr_array([plain]), 1, None)
    torch.cuda.synchronize()
    assert torch.equal(dst, plain)
    got = rec.double().cpu().numpy().reshape((-1, 4))
    assert np.isfinite(got).all(), "every node writes its record"
    g = R.stream(f64, R.Case(**kw, force_field=ff))
    rho, u = R.moments(g, R.Case(**kw, force_field=ff))
    e_rho = np.abs(got[:, 0] - rho.reshape(-1)).max()
    e_u = max(np.abs(got[:, 1 + a] - u[a].reshape(-1)).max() for a in range(case.dim))
    if case.dim == 2:
        assert (got[:, 3] == 0.0).all()
    print(f"{name} {storage}: record rho err {e_rho:.1e}, u err {e_u:.1e}")
    # measured: rho <= 2.4e-7, u <= 1.2e-8 (fp32 sums of O(1e-2) deviations)
    assert e_rho < 1e-6 and e_u < 1e-7


@pytest.mark.parametrize("lam_storage", ["f32", "f16"])
@pytest.mark.parametrize("name", list(CASES))
def test_record_adjoint_step_vs_float64_vjp_and_state_step(name, lam_storage):
    """One record-mode adjoint step: lam and dF against the torch float64 VJP of the
    restated step, and against the state-mode step on the same inputs (the same algebra on
    rho, u computed by the forward's collision instead of the adjoint's own sums: BGK sums
    in the same order, so it gives the same bits; TRT sums over opposite pairs)."""
    from pylbm_amd import _lib
    from pylbm_amd.adjoint import _LAM_DTYPES, RolloutConfig, _Engine

    kw, case, ff, f64, fs, p, eng, rng = _setup(name, "f32s", 12)
    q, d = case.lat.q, case.dim
    eng = _Engine(RolloutConfig(p, 1, storage="f32s", lam_storage=lam_storage),
                  _dev(ff.astype(np.float32)).reshape(d, -1), case.tau)
    lt = _LAM_DTYPES[lam_storage]
    lam = (0.05 * rng.standard_normal((q,) + case.array_shape)).astype(np.float32)
    lam = torch.as_tensor(lam).to(lt).double().numpy()  # representable in lam_storage
    _, rec = _record(eng, fs)
    out = {}
    for mode in ("record", "state"):
        lin = _dev(lam).to(lt).reshape(-1)
        lout = torch.full_like(lin, float("nan"))
        dF = torch.zeros(d * p.n_nodes, dtype=torch.float32, device="cuda")
        if mode == "record":
            _lib.call("pylbm_hip_adjoint_step_record", eng.ref, rec.data_ptr(),
                      lin.data_ptr(), lout.data_ptr(), dF.data_ptr(), None)
        else:
            _lib.call("pylbm_hip_adjoint_step", eng.ref, fs.data_ptr(), lin.data_ptr(),
                      lout.data_ptr(), dF.data_ptr(), None, None)
        torch.cuda.synchronize()
        out[mode] = (lout.double().cpu().numpy(), dF.double().cpu().numpy())
    assert np.isfinite(out["record"][0]).all(), "every adjoint slot is written once"
    tl = TorchLBM(R.Case(**kw, force_field=ff))
    ft = torch.tensor(f64, requires_grad=True)
    Ft = torch.tensor(ff, requires_grad=True)
    o = tl.step(ft, Ft, torch.tensor(case.tau, dtype=torch.float64))
    gf, gF = torch.autograd.grad((o * torch.tensor(lam)).sum(), (ft, Ft))
    rel = lambda a, b: np.abs(a - b).max() / np.abs(b).max()  # noqa: E731
    e_f = rel(out["record"][0].reshape(gf.shape), gf.numpy())
    e_F = rel(out["record"][1].reshape(gF.shape), gF.numpy())
    e_s = rel(out["record"][0], out["state"][0])
    e_sF = rel(out["record"][1], out["state"][1])
    same = np.array_equal(out["record"][0], out["state"][0])
    print(f"{name} lam {lam_storage}: vs f64 lam {e_f:.1e} dF {e_F:.1e}; vs state step "
          f"lam {e_s:.1e} dF {e_sF:.1e} (bit-equal: {same})")
    if lam_storage == "f32":
        # measured: lam <= 2e-7, dF <= 2e-7 (the state-mode step's figures)
        assert e_f < 2e-6 and e_F < 2e-6
    else:
        # lam_out rounded to fp16: half an fp16 ulp relative at most
        assert e_f < 1e-3 and e_F < 2e-6
    # both modes compute in fp32 from the same values: fp32 rounding apart at most (one fp16
    # ulp where a result sits at an fp16 rounding tie)
    assert e_s < (1e-6 if lam_storage == "f32" else 1e-3) and e_sF < 1e-6
    if kw.get("collision", "bgk") == "bgk":
        assert same and np.array_equal(out["record"][1], out["state"][1])


ADJ_CASES = [((40, 28), (False, False), 6, 0.05, "f32", "trt"),
             ((300, 40), (False, True), 5, 0.0, "f32", "bgk"),
             ((20, 12, 10), (False, True, False), 6, 0.04, "f32", "trt"),
             ((32, 32, 32), (True, True, True), 5, 0.0, "f32", "bgk"),
             ((32, 32, 32), (True, True, True), 6, 0.0, "f16", "bgk"),
             ((40, 28), (False, False), 5, 0.05, "f16", "trt")]


@pytest.mark.parametrize("shape,periodic,n,u_lid,lam_storage,coll", ADJ_CASES)
def test_record_adjoint_run_equals_single_record_steps(shape, periodic, n, u_lid, lam_storage,
                                                       coll):
    """pylbm_hip_adjoint_run_record (in place in pairs, alternating sweeps, an odd last step
    two-buffer) equals single two-buffer record steps, bit for bit, lam and dF; the records
    come from pylbm_hip_run_record over the same steps, whose states equal
    pylbm_hip_run_store's."""
    from pylbm_amd import _lib
    from pylbm_amd.adjoint import _LAM_DTYPES, RolloutConfig, _Engine, _ptr_array

    rng = np.random.default_rng(9)
    d = len(shape)
    ashape = tuple(reversed(shape))
    ff = (1e-5 * rng.standard_normal((d,) + ashape)).astype(np.float32)
    p = Problem(shape, tau=0.7, periodic=periodic, collision=coll, u_lid=u_lid,
                force_field=ff.astype(np.float64))
    q, N = p.lattice.q, p.n_nodes
    ffd = torch.as_tensor(ff, device="cuda").reshape(d, -1)
    eng = _Engine(RolloutConfig(p, 1, storage="f32s", lam_storage=lam_storage), ffd, 0.7)
    f0 = eng.state()
    u0 = torch.as_tensor(0.02 * rng.standard_normal((d,) + ashape), device="cuda")
    eng.init(None, u0, f0)
    states = [eng.state() for _ in range(n)]
    recs = [eng.record() for _ in range(n)]
    _lib.call("pylbm_hip_run_record", eng.ref, f0.data_ptr(), _ptr_array(states),
              _ptr_array(recs), n, None)
    plain = [eng.state() for _ in range(n)]
    _lib.call("pylbm_hip_run_store", eng.ref, f0.data_ptr(), _ptr_array(plain), n, None)
    lam0 = torch.as_tensor(0.01 * rng.standard_normal(q * N), dtype=torch.float32,
                           device="cuda").to(_LAM_DTYPES[lam_storage])
    out = []
    for mode in ("run", "steps"):
        lam = [lam0.clone(), torch.full_like(lam0, float("nan"))]
        dF = torch.zeros(d * N, dtype=torch.float32, device="cuda")
        if mode == "run":
            which = ctypes.c_int(0)
            _lib.call("pylbm_hip_adjoint_run_record", eng.ref, _ptr_array(recs), n,
                      lam[0].data_ptr(), lam[1].data_ptr(), dF.data_ptr(),
                      ctypes.byref(which), None)
            res = lam[which.value]
        else:
            cur = 0
            for k in range(n - 1, -1, -1):
                _lib.call("pylbm_hip_adjoint_step_record", eng.ref, recs[k].data_ptr(),
                          lam[cur].data_ptr(), lam[1 - cur].data_ptr(), dF.data_ptr(), None)
                cur = 1 - cur
            res = lam[cur]
        torch.cuda.synchronize()
        out.append((res.cpu(), dF.cpu()))
    for a, b in zip(states, plain):
        assert torch.equal(a, b)
    assert torch.isfinite(out[0][0].float()).all()
    assert torch.equal(out[0][0], out[1][0])
    assert torch.equal(out[0][1], out[1][1])


def test_rollout_with_records_vs_stored_states_and_float64():
    """D3Q19 BGK periodic 32^3, tau = 0.7, T = 100, C = 25, fp16 forward storage: dL/du0
    with moment records (the default when tau is not differentiated) against the stored-
    state adjoint (the same bits: BGK records hold the adjoint's own sums) and against
    float64 autograd."""
    from pylbm_amd.adjoint import RolloutConfig, rollout, velocity_loss

    n, T = 32, 100
    u0 = divergence_free_field(n, 8, 0.02, 1)
    uo = divergence_free_field(n, 8, 0.02, 2)
    kw = dict(shape=(n, n, n), tau=0.7, periodic=(True, True, True))
    grads = {}
    for records in (None, False, True):
        cfg = RolloutConfig(Problem(**kw), T, checkpoint=25, storage="f16", records=records)
        U0 = _dev(u0).requires_grad_(True)
        L = velocity_loss(rollout(cfg, u0=U0), _dev(uo))
        L.backward()
        grads[records] = (float(L), U0.grad.cpu().numpy())
    assert np.array_equal(grads[None][1], grads[True][1])
    tl = TorchLBM(R.Case(**kw))
    Ut = torch.tensor(u0, requires_grad=True)
    zero = torch.zeros(3, n, n, n, dtype=torch.float64)
    fT = tl.rollout(tl.equilibrium(torch.ones(n, n, n, dtype=torch.float64), Ut), zero,
                    torch.tensor(0.7, dtype=torch.float64), T, segment=25)
    Lt = 0.5 * ((tl.velocity(fT, zero) - torch.tensor(uo)) ** 2).sum()
    (g,) = torch.autograd.grad(Lt, (Ut,))
    g = g.numpy()
    err = np.abs(grads[True][1] - g).max() / np.abs(g).max()
    d_mode = np.abs(grads[True][1] - grads[False][1]).max() / np.abs(grads[False][1]).max()
    print(f"32^3 records: du0 rel err vs f64 {err:.2e}; records vs states {d_mode:.1e}")
    assert grads[True][0] == grads[False][0]
    assert err < 3e-4  # the fp16-forward bound of the stored-state test (measured 3.1e-5)
    assert d_mode < 1e-6


def test_channel_force_gradient_with_records():
    """D2Q9 TRT channel with a per-cell force, tau not differentiated: dL/dF with records
    against the stored-state adjoint (TRT: the records' pair sums round differently from
    the adjoint's own, within fp32 rounding) and with overlap on and off (bit for bit)."""
    from pylbm_amd.adjoint import RolloutConfig, rollout, velocity_loss

    nx, ny, T = 96, 40, 250
    rng = np.random.default_rng(7)
    ff = np.zeros((2, ny, nx))
    ff[0] = 1e-6
    ff += 1e-7 * rng.standard_normal((2, ny, nx))
    ff = ff.astype(np.float32).astype(np.float64)
    ut = smooth_random_field((nx, ny), 4, 0.01, 8)
    p = Problem((nx, ny), tau=0.75, periodic=(True, False), collision="trt", force_field=ff)
    grads = []
    for records, overlap in ((True, True), (True, False), (False, True)):
        cfg = RolloutConfig(p, T, checkpoint=40, records=records, overlap=overlap)
        F = _dev(ff).float().requires_grad_(True)
        velocity_loss(rollout(cfg, F), _dev(ut)).backward()
        grads.append(F.grad.cpu().numpy())
    assert np.isfinite(grads[0]).all()
    assert np.array_equal(grads[0], grads[1])
    d = np.abs(grads[0] - grads[2]).max() / np.abs(grads[2]).max()
    print(f"channel dL/dF records vs states: {d:.1e}")
    assert d < 1e-5


def test_records_refuse_the_tau_gradient():
    from pylbm_amd.adjoint import RolloutConfig, rollout

    p = Problem((16, 8), tau=0.8, periodic=(True, False))
    cfg = RolloutConfig(p, 4, checkpoint=2, records=True)
    tau = torch.tensor(0.8, dtype=torch.float64, device="cuda", requires_grad=True)
    with pytest.raises(ValueError, match="records"):
        rollout(cfg, tau=tau)


PAIR_CASES = [((32, 32, 32), "bgk", 6, (0.0, 0.0, 0.0)),
              ((32, 32, 32), "trt", 5, (1e-5, 0.0, 0.0)),    # odd: a last single step
              ((40, 13, 10), "bgk", 4, (0.0, 0.0, 0.0)),     # partial y-tile, empty z-chunks
              ((128, 8, 3), "trt", 2, (0.0, 2e-5, 0.0)),     # fewer planes than z-chunks
              ((100, 6, 9), "bgk", 3, (0.0, 0.0, 0.0)),      # rows shorter than 128 lanes
              ((128, 4, 1), "trt", 2, (0.0, 0.0, 0.0))]      # a single plane


@pytest.mark.parametrize("shape,coll,n,force", PAIR_CASES)
def test_pair_kernel_equals_single_record_steps(shape, coll, n, force):
    """On D3Q19 boxes periodic on every axis (fp32 adjoint states, no force field),
    pylbm_hip_adjoint_run_record takes two steps per launch, holding the first step's
    result in LDS; it equals single two-buffer record steps bit for bit."""
    from pylbm_amd import _lib
    from pylbm_amd.adjoint import RolloutConfig, _Engine, _ptr_array

    rng = np.random.default_rng(13)
    ashape = tuple(reversed(shape))
    p = Problem(shape, tau=0.7, periodic=(True, True, True), collision=coll, force=force)
    q, N = p.lattice.q, p.n_nodes
    eng = _Engine(RolloutConfig(p, 1, storage="f16"), None, 0.7)
    f0 = eng.state()
    u0 = torch.as_tensor(0.02 * rng.standard_normal((3,) + ashape), device="cuda")
    eng.init(None, u0, f0)
    states = [eng.state() for _ in range(n)]
    recs = [eng.record() for _ in range(n)]
    _lib.call("pylbm_hip_run_record", eng.ref, f0.data_ptr(), _ptr_array(states),
              _ptr_array(recs), n, None)
    lam0 = torch.as_tensor(0.01 * rng.standard_normal(q * N), dtype=torch.float32,
                           device="cuda")
    which = ctypes.c_int(0)
    lam = [lam0.clone(), torch.full_like(lam0, float("nan"))]
    _lib.call("pylbm_hip_adjoint_run_record", eng.ref, _ptr_array(recs), n, lam[0].data_ptr(),
              lam[1].data_ptr(), None, ctypes.byref(which), None)
    got = lam[which.value]
    ref = [lam0.clone(), torch.full_like(lam0, float("nan"))]
    cur = 0
    for k in range(n - 1, -1, -1):
        ref[1 - cur].fill_(float("nan"))
        _lib.call("pylbm_hip_adjoint_step_record", eng.ref, recs[k].data_ptr(),
                  ref[cur].data_ptr(), ref[1 - cur].data_ptr(), None, None)
        cur = 1 - cur
    torch.cuda.synchronize()
    assert torch.isfinite(got).all(), "every adjoint slot is written once"
    assert torch.equal(got, ref[cur])
