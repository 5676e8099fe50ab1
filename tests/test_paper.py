"""The paper's numbers and figures on the CPU.

The reference generated its paper/results.tex from three records it ships (the benchmark
record results/rtx3070-laptop.json, results/validation.json and results/counterparts.json);
tests/golden/results_reference.json holds those records, trimmed, and every macro of that
results.tex with its body.  pylbm_amd.paper must write the same macros from the same
records, byte for byte; the figures must draw from them and from a record of this package.
"""

from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

import pytest

from pylbm_amd import paper, report

ROOT = Path(__file__).resolve().parents[1]
FIX = json.loads((ROOT / "tests" / "golden" / "results_reference.json").read_text())
BENCH, VAL, CP = FIX["bench_paper"], FIX["validation"], FIX["counterparts"]
TEX = FIX["results_tex"]


def test_reference_records_give_the_reference_macros():
    got = paper.macros(BENCH, VAL, CP, source=TEX["source"])
    want = TEX["macros"]
    assert sorted(got) == sorted(want)
    diff = {k: (got[k], want[k]) for k in want if got[k] != want[k]}
    assert not diff
    assert len(want) == 182


def test_results_tex_round_trips_and_keeps_the_flags():
    text = paper.results_tex(BENCH, VAL, CP, source=TEX["source"])
    parsed = paper.parse_results_tex(text)
    assert parsed["macros"] == TEX["macros"]
    assert parsed["provisional"] is TEX["provisional"] is False
    assert parsed["cpprovisional"] is TEX["cpprovisional"] is False
    # a provisional record flips the flag; a missing record counts as provisional
    prov = paper.parse_results_tex(paper.results_tex({**BENCH, "provisional": True}, None))
    assert prov["provisional"] is True and prov["cpprovisional"] is True


def test_ladder_rows_follow_the_reference_size_rule():
    lad = paper.paper_ladder(BENCH, "D2Q9")
    # the reference size is the largest lattice of the fused Triton kernel ...
    ref = max(paper.nodes(r) for r in BENCH["runs"] if r["lattice"] == "D2Q9"
              and r["collision"] == "bgk" and (r["backend"], r["variant"]) == ("triton", "fused"))
    assert ref == 4096 * 4096
    for key, r in lad.items():
        sizes = {paper.nodes(x) for x in BENCH["runs"] if (x["backend"], x["variant"]) == key
                 and x["lattice"] == "D2Q9" and x["collision"] == "bgk"}
        # ... and every rung is shown there, or at its largest size if it never reached it
        assert paper.nodes(r) == (ref if ref in sizes else max(sizes))
    rows = TEX["macros"]["ladderRows"].split("\n")
    d2_rows = [x for x in rows if x.startswith("D2Q9 & ")]
    assert [x.split(" & ")[1] for x in d2_rows] == [paper.label(k) for k in lad]


def _hip_line(copy_bw: float) -> dict:
    """A bench.py line of this package: two D2Q9 rungs and one D3Q19 rung."""
    def v(variant, lattice, size, mlups, bpu):
        gbs = mlups * bpu / 1e3
        return {"variant": variant, "lattice": lattice, "collision": "bgk", "size": size,
                "MLUPS": mlups, "bytes_per_update": bpu, "GBs": gbs, "steps": 10}
    return {"metric": "MLUPS", "device": "AMD Instinct MI355X",
            "copy_ceiling": {"copy_bandwidth_GBs": copy_bw}, "roofline": {"peak": 8000.0},
            "variants": [v("fused", "D2Q9", [4096, 4096], 90000.0, 72),
                         v("fused-fp16", "D2Q9", [4096, 4096], 180000.0, 36),
                         v("fused", "D3Q19", [256, 256, 256], 36000.0, 152)]}


def test_hip_record_macros():
    bench = report.from_bench_line(_hip_line(7200.0))
    bench["machine"]["date"] = "2026-12-04T10:00:00+00:00"
    m = paper.bench_macros(bench, "bench_line.json")
    # ceilings at 7200 GB/s: 7200e9 / 72 B = 1e11 updates/s, and so on
    assert m["ceilFpThirtyTwo"] == "100{,}000"
    assert m["ceilSixteen"] == "200{,}000"
    assert m["ceilThreeD"] == "47{,}368"
    assert m["copyBW"] == "7{,}200" and m["specBW"] == "8{,}000"
    assert m["benchSource"] == r"bench\_line.json" and m["benchDate"] == "2026-12-04"
    assert m["bestFpThirtyTwoName"] == "HIP fused" and m["bestFpThirtyTwoMLUPS"] == "90{,}000"
    assert m["bestFpThirtyTwoPct"] == "90"  # 6480 GB/s of 7200
    assert m["bestSixteenName"] == "HIP fused fp16" and m["bestSixteenPct"] == "90"
    assert m["bestName"] == "HIP fused fp16" and m["bestPctSpec"] == "81"
    assert m["bestThreeDName"] == "HIP fused" and m["bestThreeDMLUPS"] == "36{,}000"
    assert m["hipFusedMLUPS"] == "90{,}000" and m["hipHalfPct"] == "90"
    assert m["ladderRows"].count(r"\\") == 3 and m["ladderRows"].count(r"\midrule") == 1


def test_grouping():
    assert paper.grouped(1234567.8) == "1{,}234{,}568"
    assert paper.grouped(999.4) == "999"
    assert paper.grouped(1234.5678, 2) == "1{,}234.57"


@pytest.fixture(scope="module")
def figure_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("figures")


def test_figures_from_the_reference_records(figure_dir):
    pytest.importorskip("matplotlib")
    import numpy as np

    vort = {"vorticity": np.linspace(-3, 3, 40 * 200).reshape(40, 200),
            "solid": np.zeros((40, 200), dtype=bool)}
    paths = paper.figures(figure_dir, BENCH, VAL, CP, vort, formats=("png",))
    names = sorted(p.stem for p in paths)
    assert names == sorted(["ladder", "scaling", "ghia", "viscosity", "poiseuille", "karman",
                            "counterparts"])
    for p in paths:
        assert p.read_bytes()[:8] == b"\x89PNG\r\n\x1a\n" and p.stat().st_size > 5000


def test_cli_writes_tex_and_figures(tmp_path):
    pytest.importorskip("matplotlib")
    bench = tmp_path / "rtx3070-laptop.json"
    bench.write_text(json.dumps(BENCH))
    val = tmp_path / "validation.json"
    val.write_text(json.dumps(VAL))
    cp = tmp_path / "counterparts.json"
    cp.write_text(json.dumps(CP))
    tex = tmp_path / "results.tex"
    out = subprocess.run([sys.executable, "-m", "pylbm_amd", "report", "--bench", str(bench),
                          "--validation", str(val), "--counterparts", str(cp),
                          "--tex", str(tex), "--figures", str(tmp_path / "fig")],
                         cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert paper.parse_results_tex(tex.read_text())["macros"] == TEX["macros"]
    assert (tmp_path / "fig" / "ladder.pdf").exists()
    assert (tmp_path / "fig" / "poiseuille.png").exists()
