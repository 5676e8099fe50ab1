"""Results pipeline on the CPU: the report reads the reference's own result records.

``tests/golden/results_reference.json`` holds the reference's results/quick.json and
results/validation.json with long arrays cut to three entries (make_results_fixture.py).
The report must render them, pick the rungs and numbers they hold, and refresh marked
blocks of a Markdown file; ``from_bench_line`` maps the one-line bench.py record onto the
same schema.
"""

from __future__ import annotations

import json
import math
from pathlib import Path

import pytest

from pylbm_amd import report

FIX = json.loads((Path(__file__).parent / "golden" / "results_reference.json").read_text())
BENCH, VAL = FIX["bench"], FIX["validation"]


def test_ladder_takes_each_rungs_largest_size():
    rungs = report.ladder(BENCH, "D2Q9")
    keys = [report.rung_key(r) for r in rungs]
    assert len(keys) == len(set(keys))
    expect = {}
    for r in BENCH["runs"]:
        if r["lattice"] == "D2Q9" and r["collision"] == "bgk":
            n = math.prod(r["size"])
            k = (r["backend"], r["variant"])
            expect[k] = max(expect.get(k, 0), n)
    assert set(keys) == set(expect)
    for r in rungs:
        assert math.prod(r["size"]) == expect[report.rung_key(r)]


def test_headline_numbers_come_from_the_record():
    d2 = [r for r in report.ladder(BENCH, "D2Q9") if r["backend"] != "numpy"]
    pull = [r for r in d2 if "-aa" not in r["variant"]]
    aa = [r for r in d2 if "-aa" in r["variant"]]
    best32 = max((r for r in pull if "16" not in r["variant"]), key=lambda r: r["median"])
    best16 = max((r for r in pull if "16" in r["variant"]), key=lambda r: r["median"])
    aa32 = max((r for r in aa if "16" not in r["variant"]), key=lambda r: r["median"])
    lines = report.headline(BENCH)
    assert f"{best32['median']:,.0f} MLUPS" in lines[0]
    assert f"{best32['backend']}/{best32['variant']}" in lines[0]
    assert f"{best16['median']:,.0f} MLUPS" in lines[1]
    assert f"{best16['median'] / best32['median']:.2f}x" in lines[1]
    # the AA kernels are reported on their own line, marked as cache-assisted
    assert "cache-assisted" in lines[2] and f"{aa32['median']:,.0f} MLUPS" in lines[2]
    # the float64 NumPy rung is never the headline
    assert "numpy" not in " ".join(lines)


def test_performance_table_has_a_row_per_rung():
    table = report.performance_table(BENCH)
    rows = [ln for ln in table.splitlines() if ln.startswith("| D")]
    n = sum(len(report.ladder(BENCH, lat, c)) for lat in ("D2Q9", "D3Q19")
            for c in ("bgk", "trt", "mrt"))
    assert len(rows) == n
    bw = BENCH["ceiling"]["copy_bandwidth_GBs"]
    assert f"{bw * 1e9 / 72 / 1e6:,.0f}" in table


def test_validation_table_covers_every_block():
    table = report.validation_table(VAL)
    assert table.count("Ghia et al. 1982") == len(VAL["ghia"])
    worst = max(r["rel_error"] for r in VAL["shear_wave"])
    assert f"{100 * worst:.2f}%" in table
    assert f"order {VAL['poiseuille']['bgk']['order']:.3f}" in table
    assert f"St = {VAL['karman']['St']:.4f}" in table
    for b in {r["backend"] for r in VAL["taylor_green_3d"]}:
        assert f"({b})" in table
    assert "equivalence, f16 storage" in table


def test_fill_blocks_replaces_only_named_blocks():
    text = ("# T\n<!-- pylbm_amd:performance -->\nold\n<!-- /pylbm_amd:performance -->\n"
            "<!-- pylbm_amd:validation -->\nkeep\n<!-- /pylbm_amd:validation -->\n")
    out = report.fill_blocks(text, {"performance": "new table"})
    assert "new table" in out and "old" not in out and "keep" in out
    assert out.count("<!-- pylbm_amd:performance -->") == 1
    again = report.fill_blocks(out, {"performance": "newer"})
    assert "newer" in again and "new table" not in again
    with pytest.raises(KeyError):
        report.fill_blocks(text, {"headline": "x"})


def test_bench_line_converts_to_the_schema(tmp_path):
    line = {"metric": "MLUPS", "value": 80000.0, "device": "MI355X",
            "roofline": {"peak": 8000.0},
            "copy_ceiling": {"copy_bandwidth_GBs": 6000.0},
            "variants": [
                {"variant": "fused", "collision": "bgk", "lattice": "D2Q9", "size": [64, 64],
                 "storage": "f32", "steps": 10, "MLUPS": 80000.0, "GBs": 5760.0,
                 "bytes_per_update": 72},
                {"variant": "gather", "collision": "trt", "lattice": "D2Q9", "size": [64, 64],
                 "storage": "f32", "steps": 10, "MLUPS": 70000.0, "GBs": 5040.0,
                 "bytes_per_update": 72, "workload": "channel"}]}
    p = tmp_path / "bench.json"
    p.write_text("log line\n" + json.dumps(line) + "\n")
    rec = report.load(p)
    keys = set(BENCH["runs"][0]) - {"dtype", "notes", "memory_MiB", "flops_per_unit",
                                     "pct_peak_flops"}
    for r in rec["runs"]:
        assert keys <= set(r)
    assert rec["runs"][0]["pct_copy_bw"] == pytest.approx(96.0)
    assert [report.rung_key(r) for r in report.ladder(rec, "D2Q9", "trt")] == []
    assert len(report.ladder(rec, "D2Q9", "trt", workload=True)) == 1
    assert "80,000 MLUPS" in report.render(rec)
    # the metric's own number leads the headline; its duplicate leg stays out of the table
    line["variants"].append(dict(line["variants"][0], MLUPS=79000.0, GBs=5688.0,
                                 protocol="headline"))
    line["config"] = {"workload": "D2Q9 4096^2"}
    rec = report.from_bench_line(line)
    assert len(rec["runs"]) == 2 and rec["headline"]["value"] == 80000.0
    assert report.headline(rec)[0].startswith("bench.py on MI355X: **80,000 MLUPS**")


def test_cli_report_writes_markdown(tmp_path):
    from pylbm_amd.cli import main

    b = tmp_path / "b.json"
    v = tmp_path / "v.json"
    b.write_text(json.dumps(BENCH))
    v.write_text(json.dumps(VAL))
    readme = tmp_path / "R.md"
    readme.write_text("<!-- pylbm_amd:headline -->\n<!-- /pylbm_amd:headline -->\n"
                      "<!-- pylbm_amd:performance -->\n<!-- /pylbm_amd:performance -->\n"
                      "<!-- pylbm_amd:validation -->\n<!-- /pylbm_amd:validation -->\n")
    out = tmp_path / "r.md"
    main(["report", "--bench", str(b), "--validation", str(v), "--out", str(out),
          "--readme", str(readme)])
    text = out.read_text()
    assert "## Performance" in text and "## Validation" in text
    assert "| D2Q9 | bgk |" in readme.read_text()
