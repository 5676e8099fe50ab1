"""Result records written on the GPU carry the reference's schema (GPU).

The quick validation suite and one benchmark measurement run on the hip backend; every
key of the matching record in the reference's published results (the trimmed fixture
tests/golden/results_reference.json) must be present, the physics must pass loose bounds,
and the report must render both records.
"""

from __future__ import annotations

import json
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

FIX = json.loads((Path(__file__).parent / "golden" / "results_reference.json").read_text())
REF_VAL, REF_BENCH = FIX["validation"], FIX["bench"]


@pytest.fixture(scope="module")
def suite(tmp_path_factory):
    from pylbm_amd.validation.suite import run_suite

    out = tmp_path_factory.mktemp("val") / "validation.json"
    lines = []
    rec = run_suite(quick=True, out=out, log=lines.append)
    print("\n".join(lines), flush=True)
    assert json.loads(out.read_text()) == json.loads(json.dumps(rec))
    return rec


def test_suite_blocks_have_the_reference_keys(suite):
    assert set(REF_VAL) - {"equivalence"} <= set(suite)
    for block in ("ghia", "shear_wave", "taylor_green_3d"):
        assert suite[block], block
        for r in suite[block]:
            assert set(REF_VAL[block][0]) <= set(r), block
    assert set(REF_VAL["ghia"][0]["errors"]) <= set(suite["ghia"][0]["errors"])
    assert set(REF_VAL["ghia"][0]["profile"]) <= set(suite["ghia"][0]["profile"])
    for coll in ("bgk", "trt"):
        assert set(REF_VAL["poiseuille"][coll]) <= set(suite["poiseuille"][coll])
        for r in suite["poiseuille"][coll]["runs"]:
            assert set(REF_VAL["poiseuille"][coll]["runs"][0]) <= set(r)
    assert set(REF_VAL["karman"]) <= set(suite["karman"])


def test_suite_physics_within_loose_bounds(suite):
    assert all(r["backend"].startswith("hip/") for r in suite["shear_wave"])
    assert max(r["rel_error"] for r in suite["shear_wave"]) < 2e-2
    assert max(r["rel_error"] for r in suite["taylor_green_3d"]) < 5e-2
    assert all(r["backend"] == "hip/fused-f32s" for r in suite["poiseuille"]["bgk"]["runs"])
    assert suite["poiseuille"]["bgk"]["order"] == pytest.approx(2.0, abs=0.1)
    # TRT is exact in float64 (test_validation.py); what remains is the fp32 rounding
    # offset of the steady state (2e-5 at h = 16 on shifted fp32 storage; the reference's
    # own plain-fp32 run reports L2 = 1.1e-3 at h = 32)
    assert max(r["L2"] for r in suite["poiseuille"]["trt"]["runs"]) < 1e-4
    g = suite["ghia"][0]
    assert g["converged"] and g["errors"]["Linf"] < 0.02
    k = suite["karman"]
    assert 0.25 < k["St"] < 0.35 and 2.5 < k["cd_max"] < 4.0


def test_bench_record_has_the_reference_keys(suite):
    from pylbm_amd import report
    from pylbm_amd.benchmark import measure

    r = measure("fused", 2, 64, {"min_time": 0.02, "repeats": 2}, copy_bw=5000.0,
                log=lambda s: None)
    assert set(REF_BENCH["runs"][0]) <= set(r)
    assert r["backend"] == "hip" and r["median"] > 0 and r["size"] == [64, 64]
    bench = {"machine": {"gpu": "test"}, "ceiling": {"copy_bandwidth_GBs": 5000.0,
                                                     "spec_bandwidth_GBs": 8000.0},
             "runs": [r]}
    text = report.render(bench, suite)
    assert "hip/fused" in text and "Schäfer–Turek" in text
