"""The command line (pylbm/cli.py:98-144 names): ``backends`` everywhere, ``run`` on the GPU.

``backends`` must list every registry entry and never raise, also without a GPU
(reference tests/test_codegen_cli.py:24-29).  ``run cavity --backend hip --variant V`` is the
entry the hip backend is reached by; on the GPU each of the six fused variants runs a small
cavity through it, and the state it leaves (rho, u saved with ``--out``) is compared with
the float64 oracle after the same number of steps.
"""

from __future__ import annotations

import json

import numpy as np
import pytest

from oracle import lbm_ref as R
from pylbm_amd import backends as B
from pylbm_amd.cli import main, parser
from pylbm_amd.problem import Problem

NORTH_STAR = ["fused", "fused-aa", "fused-fp16", "fused-bf16", "fused-aa-fp16",
              "fused-aa-bf16"]
#: max |u - u_oracle| after ~200 steps of a 48^2 cavity (f32 measured ~1e-7)
U_TOL = {"f32": 1e-5, "f16": 3e-4, "bf16": 3e-3}
#: relative mass change over the run (bf16 measured 1.3e-5 to 1.8e-5)
MASS_TOL = {"f32": 1e-6, "f16": 1e-5, "bf16": 1e-4}


def test_backends_lists_every_entry_and_never_raises(capsys):
    assert main(["backends"]) == 0
    out = capsys.readouterr().out
    for backend, variant in B.entries():
        assert f"{backend}/{variant}" in out
    assert "AA in place" in out and "two buffers" in out


def test_run_flags_keep_the_reference_names():
    a = parser().parse_args(["run", "cavity", "--n", "64", "--re", "400", "--u-lid", "0.05",
                             "--collision", "trt", "--backend", "hip", "--variant",
                             "fused-aa-fp16", "--steps", "100", "--out", "x.npz"])
    assert (a.case, a.n, a.re, a.u_lid, a.collision, a.backend, a.variant, a.steps, a.out) == \
        ("cavity", 64, 400.0, 0.05, "trt", "hip", "fused-aa-fp16", 100, "x.npz")
    for cmd in (["validate", "--quick", "--cases", "ghia"], ["bench", "--dev"],
                ["report", "--bench", "b.json"]):
        parser().parse_args(cmd)


def test_run_without_a_gpu_fails_with_the_reason():
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(RuntimeError, match="unavailable"):
        main(["run", "cavity", "--n", "16", "--steps", "2", "--backend", "hip"])


@pytest.mark.gpu
@pytest.mark.parametrize("variant", NORTH_STAR)
def test_run_cavity_north_star_variants_match_oracle(variant, tmp_path, capsys):
    n, re, u_lid = 48, 100.0, 0.1
    out = tmp_path / "state.npz"
    assert main(["run", "cavity", "--n", str(n), "--re", str(re), "--u-lid", str(u_lid),
                 "--steps", "200", "--backend", "hip", "--variant", variant, "--json",
                 "--out", str(out)]) == 0
    summary = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    assert summary["entry"] == f"hip/{variant}"
    assert summary["steps"] == 200 and summary["MLUPS"] > 0
    # 16-bit storage rounds every stored value, so mass drifts by its rounding only
    assert abs(summary["mass_change"]) < MASS_TOL[summary["storage"]]
    saved = np.load(out)
    steps = int(saved["steps"])
    assert steps == summary["steps_total"] == 200 + (2 if B.get("hip", variant).aa else 1)
    p = Problem.cavity(n, re=re, u_lid=u_lid)
    case = R.Case(shape=(n, n), tau=p.tau, u_lid=u_lid)
    rho, u = R.macroscopic(R.run(case, steps), case)
    err_u = float(np.abs(saved["u"] - u).max())
    err_rho = float(np.abs(saved["rho"] - rho).max())
    print(f"run cavity hip/{variant}: {steps} steps, |du| {err_u:.2e}, |drho| {err_rho:.2e}")
    tol = U_TOL[summary["storage"]]
    assert err_u < tol and err_rho < 3 * tol
    assert abs(summary["max_speed"] - float(np.sqrt((u ** 2).sum(0)).max())) < tol


@pytest.mark.gpu
def test_run_cavity3d_and_karman(tmp_path, capsys):
    assert main(["run", "cavity3d", "--n", "16", "--re", "50", "--collision", "trt",
                 "--steps", "40", "--backend", "hip", "--variant", "fused-aa", "--json",
                 "--out", str(tmp_path / "c3.npz")]) == 0
    s3 = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    p = Problem.cavity(16, re=50.0, u_lid=0.1, collision="trt", dim=3)
    case = R.Case(shape=(16, 16, 16), tau=p.tau, u_lid=0.1, collision="trt")
    _, u = R.macroscopic(R.run(case, s3["steps_total"]), case)
    assert float(np.abs(np.load(tmp_path / "c3.npz")["u"] - u).max()) < 1e-5
    # the Schaefer-Turek channel (obstacle, inlet, outlet) with the default entry
    assert main(["run", "karman", "--d", "10", "--steps", "100", "--backend", "hip",
                 "--json"]) == 0
    sk = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    assert sk["entry"] == "hip/fused" and sk["lattice"] == "D2Q9"
    assert np.isfinite(sk["mass"]) and 0.0 < sk["max_speed"] < 0.2
