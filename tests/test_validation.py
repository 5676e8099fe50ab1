"""Analytic validation (mirrors the reference's tests/test_validation.py).

CPU tests run the validation cases on the float64 oracle (pinning both the oracle and the
measurement code); GPU tests run them on the hip backend with the reference's GPU bounds.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import lbm_ref as R
from pylbm_amd.validation import ghia, poiseuille, shear_wave, taylor_green


def oracle_factory(p):
    return R.OracleSolver(R.Case(
        shape=p.shape, tau=p.tau, collision=p.collision, periodic=p.periodic, u_lid=p.u_lid,
        force=p.force, magic=p.magic, s_e=p.s_e, s_eps=p.s_eps, s_q=p.s_q,
        force_field=p.force_field, init_rho=p.init_rho, init_u=p.init_u))


@pytest.mark.parametrize("tau", [0.6, 1.0, 1.4])
def test_shear_wave_viscosity_cpu(tau):
    r = shear_wave.run_shear_wave(tau, n=32, factory=oracle_factory)
    assert r["rel_error"] < 0.02


def test_poiseuille_bgk_second_order_and_trt_exact_cpu():
    c = poiseuille.convergence((8, 16), collision="bgk", factory=oracle_factory)
    assert 1.9 < c["order"] < 2.1
    r = poiseuille.run_poiseuille(12, collision="trt", factory=oracle_factory)
    assert r["L2"] < 1e-9


def test_tgv3d_viscosity_cpu():
    r = taylor_green.run_taylor_green(n=16, plane="yz", factory=oracle_factory)
    assert r["rel_error"] < 0.03


def test_ghia_table_and_exclusion():
    ref = ghia.reference(400)
    assert len(ref["y"]) == len(ref["u"]) == 17
    assert ("v", "400", 0.9063) in ghia.EXCLUDE
    # the centreline helper puts node centres at (i + 1/2)/n and adds the wall values
    prof = ghia.centerlines(np.zeros((4, 4)), np.zeros((4, 4)), 0.1)
    assert np.allclose(prof["y"], [0, 0.125, 0.375, 0.625, 0.875, 1])
    assert prof["u"][-1] == 1.0 and prof["v"][-1] == 0.0


# ----------------------------------------------------------------------------- GPU


@pytest.mark.gpu
def test_shear_wave_gpu():
    r = shear_wave.run_shear_wave(0.8, n=128, backend="hip", variant="fused")
    assert r["rel_error"] < 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["fused", "fused-aa", "fused-fp16"])
def test_tgv3d_gpu(variant):
    r = taylor_green.run_taylor_green(n=48, plane="xz", backend="hip", variant=variant)
    assert r["rel_error"] < 5e-3


@pytest.mark.gpu
def test_poiseuille_trt_gpu_fp32():
    """fp32 round-off floor: the Guo increment (~5e-6) is resolved to ~1e-3 relative in
    populations of ~0.1, so the steady profile sits ~1e-3 off the parabola; the reference's
    own fp32 run reports L2 = 1.1e-3 at h = 32 (results/validation.json, bgk_fp32_h32)."""
    r = poiseuille.run_poiseuille(32, collision="trt", backend="hip", variant="fused",
                                  tol=1e-7)
    print(f"Poiseuille TRT fp32 h=32: L2 = {r['L2']:.2e} after {r['steps']} steps")
    assert r["L2"] < 2e-3


@pytest.mark.gpu
def test_ghia_re100_gpu():
    r = ghia.run_cavity(100, 128, backend="hip", variant="fused")
    assert r.converged
    assert r.errors["Linf"] < 0.006


#: The reference's published Ghia runs (results/validation.json, "ghia"): Re, n, collision
#: -> (steps to convergence, L2, Linf), all fp32 two-buffer fused kernels on its GPU.
GHIA_PUBLISHED = {
    (1000, 256, "bgk"): (129024, 0.00400, 0.00805),
    (1000, 256, "trt"): (131072, 0.00404, 0.00742),
    (1000, 256, "mrt"): (135168, 0.00395, 0.00823),
    (400, 256, "bgk"): (87040, 0.00237, 0.00470),
}


@pytest.mark.gpu
@pytest.mark.parametrize("key", list(GHIA_PUBLISHED), ids=lambda k: f"re{k[0]}-{k[2]}")
def test_ghia_matches_reference_published_runs(key):
    """Same discrete problem, same steady-state test, against the reference's own fp32 runs.

    The test stops at an RMS change of 1e-5 u_lid per check while a slow transient is still
    decaying: tightening it to 1e-7 moves Linf by a further 3e-4 (BGK, TRT and MRT alike),
    and fp32 rounding differences shift the stopping check by a few checks.  So the bound
    is the L2 error within 3e-4, Linf within 1.5e-3, and the step count within 10%.  MRT
    reproduces the published run to every printed digit (135168 steps, L2 0.00395, Linf
    0.00823)."""
    re, n, coll = key
    steps, l2, linf = GHIA_PUBLISHED[key]
    r = ghia.run_cavity(re, n, collision=coll, backend="hip", variant="fused")
    print(f"Ghia Re={re} {n}^2 {coll}: steps {r.steps} (reference {steps}), "
          f"L2 {r.errors['L2']:.5f} ({l2}), Linf {r.errors['Linf']:.5f} ({linf})")
    assert r.converged is True
    assert abs(r.steps - steps) <= 0.1 * steps
    assert abs(r.errors["L2"] - l2) < 3e-4
    assert abs(r.errors["Linf"] - linf) < 1.5e-3
