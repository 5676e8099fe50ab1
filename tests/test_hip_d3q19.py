"""D3Q19 at larger sizes (GPU): periodic TRT Taylor-Green against the C float64 oracle.

Config-3 correctness: D3Q19 TRT (tau+ = 0.55, Lambda = 3/16), periodic 64^3, initial state
the equilibrium of u = 0.02 (sin x cos y cos z, -cos x sin y cos z, 0), rho = 1; after 500
steps the velocity Linf against float64 is reported for every storage format.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import lbm_ref as R
from oracle import lbm_ref_c as RC
from pylbm_amd import backends as B
from pylbm_amd.problem import Problem
from tests.cases import tgv3d_fields

pytestmark = pytest.mark.gpu

N3, STEPS = 64, 500
#: velocity Linf bounds after 500 steps (|u| <= 0.02)
BOUND = {"f32": 1e-6, "f16": 2e-4, "bf16": 2e-3}


@pytest.fixture(scope="module")
def tgv_ref():
    rho, u = tgv3d_fields(N3)
    kw = dict(shape=(N3,) * 3, tau=0.55, collision="trt", periodic=(True,) * 3, init_rho=rho,
              init_u=u)
    case = R.Case(**kw)
    f = RC.run(case, STEPS)
    return kw, f, R.macroscopic(f, case)[1]


@pytest.mark.parametrize("variant", ["fused", "fused-fp16", "fused-bf16", "fused-aa",
                                     "fused-aa-fp16", "fused-aa-bf16"])
def test_tgv64_500_steps_velocity_linf(variant, tgv_ref):
    kw, f_ref, u_ref = tgv_ref
    s = B.make("hip", variant, Problem(**kw))
    s.step(STEPS)
    _, u = s.macroscopic()
    linf = float(np.abs(u - u_ref).max())
    df = float(np.abs(s.populations() - f_ref).max())
    print(f"TGV 64^3 TRT 500 steps hip/{variant}: u Linf = {linf:.3e}, max|df| = {df:.3e}")
    assert linf < BOUND[s.storage]
    if s.storage == "f32":
        assert df <= 1e-5


@pytest.mark.parametrize("variant", ["fused", "fused-aa"])
def test_d3q19_walls_force_32(variant):
    """D3Q19 BGK, periodic x, walls y, lid on z, uniform force, 32 x 24 x 20, 200 steps."""
    kw = dict(shape=(32, 24, 20), tau=0.65, periodic=(True, False, False), u_lid=0.05,
              force=(1e-5, 2e-6, -1e-6))
    ref = RC.run(R.Case(**kw), 200)
    s = B.make("hip", variant, Problem(**kw))
    s.step(200)
    err = float(np.abs(s.populations() - ref).max())
    print(f"D3Q19 walls+lid+force hip/{variant}: max|df| = {err:.3e}")
    assert err <= 1e-5
