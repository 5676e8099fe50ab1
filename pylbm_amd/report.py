"""Markdown report from benchmark and validation results.

The report side of the reference's results pipeline (figures/make_figures.py:400-651,
which fills README and paper tables from results/*.json): a headline, the performance
ladder per lattice and collision, and the validation table.  It reads

- a benchmark record in the reference's schema (``machine``, ``ceiling``, ``runs`` with
  ``backend``, ``variant``, ``lattice``, ``collision``, ``size``, ``median``, ``min``,
  ``max``, ``bytes_per_unit``, ``effective_GBs``, ``pct_copy_bw``, ``pct_spec_bw``), as
  ``python -m pylbm_amd bench --out`` writes it, or the one-line JSON of ``bench.py``
  (converted by ``from_bench_line``);
- a validation record in the reference's schema, as ``python -m pylbm_amd validate --out``
  writes it.

The reference's own result files follow the same schema and render unchanged.  Blocks of a
Markdown file between ``<!-- pylbm_amd:NAME -->`` and ``<!-- /pylbm_amd:NAME -->`` can be
refreshed in place (``fill_blocks``).
"""

from __future__ import annotations

import json
import math
import re
from pathlib import Path
from typing import Any, Iterable

#: hip variants in ladder order: two-buffer before AA, fp32 before 16-bit storage, one step
#: per launch before two
HIP_ORDER = ["gather", "fused", "fused-f32s", "fused-aa", "fused-fp16", "fused-bf16",
             "fused-aa-fp16", "fused-aa-bf16", "fused-2step", "fused-2step-fp16",
             "fused-2step-bf16"]


def load(path: str | Path) -> dict:
    """A benchmark or validation record; a ``bench.py`` output file (its last non-empty
    line is the JSON record) is converted to the benchmark schema."""
    text = Path(path).read_text().strip()
    try:
        rec = json.loads(text)
    except json.JSONDecodeError:
        rec = json.loads(text.splitlines()[-1])
    if "runs" not in rec and "variants" in rec and "metric" in rec:
        rec = from_bench_line(rec)
    return rec


def from_bench_line(line: dict) -> dict:
    """The benchmark schema from one ``bench.py`` JSON line (its ``variants`` block).

    The line times the headline kernel twice: once under the metric's protocol (its
    ``value``, kept under ``headline``) and once as an ordinary leg; the table keeps the leg,
    as it keeps every other leg.  The plane-stream ceilings go to ``ceiling.plane_stream``."""
    ceil = line.get("copy_ceiling", {})
    spec = line.get("roofline", {}).get("peak", 8000.0)
    copy_bw = ceil.get("copy_bandwidth_GBs") or ceil.get("hip_copy_GBs")
    runs = []
    for v in line.get("variants", []):
        if v.get("protocol") == "headline":
            continue
        gbs = v["GBs"]
        runs.append({
            "backend": "hip", "variant": v["variant"], "lattice": v["lattice"],
            "collision": v["collision"], "size": list(v["size"]),
            "storage": v.get("storage"), "metric": "MLUPS", "value": v["MLUPS"],
            "median": v["MLUPS"], "min": v.get("MLUPS_min", v["MLUPS"]),
            "max": v.get("MLUPS_max", v["MLUPS"]), "repeats": v.get("repeats", 1),
            "steps_per_repeat": v.get("steps"),
            "bytes_per_unit": v["bytes_per_update"], "effective_GBs": gbs,
            "pct_copy_bw": 100.0 * gbs / copy_bw if copy_bw else None,
            "pct_spec_bw": 100.0 * gbs / spec,
            "workload": v.get("workload", "")})
    out = {"machine": {"gpu": line.get("device", "GPU"), "date": line.get("date", "")},
           "ceiling": {"copy_bandwidth_GBs": copy_bw, "spec_bandwidth_GBs": spec},
           "plan": "bench.py", "provisional": False, "runs": runs}
    if line.get("plane_stream"):
        out["ceiling"]["plane_stream"] = line["plane_stream"]
    if "value" in line:
        out["headline"] = {"metric": line.get("metric"), "value": line["value"],
                           "unit": line.get("unit", "MLUPS"), "steps": line.get("steps"),
                           "warmup": line.get("warmup"),
                           "workload": line.get("config", {}).get("workload", ""),
                           "pct_copy": line.get("pct_copy"),
                           "roofline_frac": line.get("roofline", {}).get("frac")}
    return out


# ----------------------------------------------------------------------------- selection
def _nodes(run: dict) -> int:
    return math.prod(int(s) for s in run["size"])


def rung_key(run: dict) -> tuple[str, str]:
    return run["backend"], run["variant"]


def _rank(key: tuple[str, str]) -> tuple[int, int, str]:
    b, v = key
    if b == "hip" and v in HIP_ORDER:
        return 0, HIP_ORDER.index(v), ""
    return 1, 0, f"{b}/{v}"


def ladder(bench: dict, lattice: str, collision: str = "bgk",
           workload: bool = False) -> list[dict]:
    """One run per (backend, variant): the largest size it measured (ties: the faster run).

    Runs that carry a non-empty ``workload`` (obstacle channels and the like) are kept out
    unless ``workload`` is True, in which case only they are taken, one run per (workload,
    backend, variant)."""
    best: dict[tuple[str, str, str], dict] = {}
    for r in bench.get("runs", []):
        if r.get("lattice") != lattice or r.get("collision") != collision:
            continue
        if bool(r.get("workload")) != workload:
            continue
        k = (r.get("workload") or "",) + rung_key(r)
        cur = best.get(k)
        if cur is None or (_nodes(r), r["median"]) > (_nodes(cur), cur["median"]):
            best[k] = r
    return [best[k] for k in sorted(best, key=lambda k: (k[0],) + _rank(k[1:]))]


def ceiling_mlups(bench: dict, bytes_per_update: float) -> float | None:
    """MLUPS at the measured copy bandwidth for a given traffic per node update."""
    bw = bench.get("ceiling", {}).get("copy_bandwidth_GBs")
    return bw * 1e9 / bytes_per_update / 1e6 if bw else None


def _is16(run: dict) -> bool:
    return "16" in run["variant"] or run.get("storage") in ("f16", "bf16")


# ------------------------------------------------------------------------------- tables
def _size(run: dict) -> str:
    return "x".join(str(s) for s in run["size"])


def _pct(x: float | None) -> str:
    return "—" if x is None else f"{x:.0f}"


def _is_aa(run: dict) -> bool:
    return "-aa" in run["variant"] or run["variant"].endswith("aa")


def _is_pair(run: dict) -> bool:
    """Two steps per launch (the hip ``fused-2step*`` rungs)."""
    return "2step" in run["variant"]


def headline(bench: dict) -> list[str]:
    """The metric's own number when the record has one, then the best D2Q9 rates: the
    two-buffer pull kernels (fp32, 16-bit and their ratio) and, beside them, the in-place AA
    kernels, whose alternating sweeps let the 256 MiB Infinity Cache serve part of each
    launch's reads (so they can pass the streaming copy), and the two-steps-per-launch
    kernels against the one-step pull; then the best fp32 D3Q19 rate.
    The reference's float64 NumPy rung is not a candidate."""
    gpu = bench.get("machine", {}).get("gpu", "GPU")
    bw = bench.get("ceiling", {}).get("copy_bandwidth_GBs")
    spec = bench.get("ceiling", {}).get("spec_bandwidth_GBs")
    out = []
    h = bench.get("headline")
    if h:
        frac = h.get("roofline_frac")
        out.append(f"bench.py on {gpu}: **{h['value']:,.0f} {h.get('unit', 'MLUPS')}** "
                   f"({h.get('workload', '')}; {h.get('warmup')} warm-up + "
                   f"{h.get('steps')} timed steps)"
                   + (f" = {h['pct_copy']:.1f}% of the measured copy bandwidth"
                      if h.get("pct_copy") is not None else "")
                   + (f", {100 * frac:.1f}% of the 8 TB/s spec" if frac else "") + ".")
    d2 = [r for r in ladder(bench, "D2Q9") if r["backend"] != "numpy"]
    pull = [r for r in d2 if not _is_aa(r) and not _is_pair(r)]
    aa = [r for r in d2 if _is_aa(r)]
    pair = [r for r in d2 if _is_pair(r)]

    def best(rs: list[dict], sixteen: bool) -> dict | None:
        c = [r for r in rs if _is16(r) == sixteen]
        return max(c, key=lambda r: r["median"]) if c else None

    p32, p16, a32, a16 = best(pull, False), best(pull, True), best(aa, False), best(aa, True)
    if p32:
        out.append(f"D2Q9 fp32, two-buffer pull: **{p32['median']:,.0f} MLUPS** "
                   f"({p32['backend']}/{p32['variant']}, {_size(p32)}) = "
                   f"{_pct(p32.get('pct_copy_bw'))}% of the measured {bw:,.0f} GB/s copy "
                   f"bandwidth, {_pct(p32.get('pct_spec_bw'))}% of the {spec:,.0f} GB/s spec.")
    if p16:
        line = (f"16-bit storage, two-buffer pull: **{p16['median']:,.0f} MLUPS** "
                f"({p16['backend']}/{p16['variant']}, {_pct(p16.get('pct_copy_bw'))}% of copy "
                f"bandwidth at {p16['bytes_per_unit']} B/update)")
        if p32:
            line += f", {p16['median'] / p32['median']:.2f}x the best fp32 pull rung"
        out.append(line + ".")
    if a32 or a16:
        parts = []
        for r, name in ((a32, "fp32"), (a16, "16-bit")):
            if r:
                parts.append(f"{name} **{r['median']:,.0f} MLUPS** ({r['backend']}/"
                             f"{r['variant']}, {_pct(r.get('pct_copy_bw'))}% of copy)")
        line = "AA in place (cache-assisted: alternating sweeps re-read from the Infinity " \
               "Cache): " + ", ".join(parts)
        if a32 and a16:
            line += f"; 16-bit / fp32 {a16['median'] / a32['median']:.2f}x"
        out.append(line + ".")
    t32, t16 = best(pair, False), best(pair, True)
    if t32 or t16:
        parts = []
        for r, one, name in ((t32, p32, "fp32"), (t16, p16, "16-bit")):
            if r:
                part = (f"{name} **{r['median']:,.0f} MLUPS** ({r['backend']}/{r['variant']}, "
                        f"{r['bytes_per_unit']} B/update")
                if one:
                    part += f", {r['median'] / one['median']:.2f}x the {name} pull rung"
                parts.append(part + ")")
        out.append("Two steps per launch (the periodic box; the state crosses memory once per "
                   "two steps, the intermediate one stays in LDS): " + ", ".join(parts) + ".")
    d3 = [r for r in ladder(bench, "D3Q19") + ladder(bench, "D3Q19", "trt")
          if r["backend"] != "numpy" and not _is16(r)]
    if d3:
        b = max(d3, key=lambda r: r["median"])
        out.append(f"D3Q19 fp32: **{b['median']:,.0f} MLUPS** ({b['backend']}/{b['variant']}, "
                   f"{b['collision'].upper()}, {_size(b)}, {_pct(b.get('pct_copy_bw'))}% of "
                   f"copy bandwidth).")
    return out


def performance_table(bench: dict, collisions: Iterable[str] = ("bgk", "trt", "mrt")) -> str:
    """Markdown ladder: one row per lattice, collision and rung, plus obstacle workloads."""
    rows = ["| lattice | collision | rung | size | MLUPS | B/update | GB/s | % copy | % spec |",
            "|---|---|---|---|---:|---:|---:|---:|---:|"]
    for lat in ("D2Q9", "D3Q19"):
        for coll in collisions:
            for wl in (False, True):
                for r in ladder(bench, lat, coll, workload=wl):
                    name = f"{r['backend']}/{r['variant']}"
                    if wl:
                        name += f" ({r['workload']})"
                    rows.append(f"| {lat} | {coll} | {name} | {_size(r)} | "
                                f"{r['median']:,.0f} | {r['bytes_per_unit']} | "
                                f"{r['effective_GBs']:,.0f} | {_pct(r.get('pct_copy_bw'))} | "
                                f"{_pct(r.get('pct_spec_bw'))} |")
    bw = bench.get("ceiling", {}).get("copy_bandwidth_GBs")
    note = (f"Ceiling: measured device copy bandwidth {bw:,.0f} GB/s"
            f" = {ceiling_mlups(bench, 72):,.0f} / {ceiling_mlups(bench, 36):,.0f} MLUPS for "
            f"D2Q9 fp32 / 16-bit and {ceiling_mlups(bench, 152):,.0f} / "
            f"{ceiling_mlups(bench, 76):,.0f} for D3Q19." if bw else "")
    ps = bench.get("ceiling", {}).get("plane_stream") or {}
    names = (("d2q9_f32", "D2Q9 fp32"), ("d2q9_16", "D2Q9 16-bit"), ("d3q19_f32", "D3Q19 fp32"),
             ("d3q19_16", "D3Q19 16-bit"))
    got = [f"{label} {ps[k + '_GBs']:,.0f}" for k, label in names if ps.get(k + "_GBs")]
    if got:
        note += (" Plane stream (the two-buffer kernels' access pattern without arithmetic, "
                 "pylbm_hip_plane_probe) at the legs' state shapes: " + ", ".join(got)
                 + " GB/s.")
    sw = [f"{label} {ps[k + '_GBs']:,.0f}" for k, label in
          (("d2q9_sweep_f32", "fp32"), ("d2q9_sweep_16", "16-bit")) if ps.get(k + "_GBs")]
    if sw:
        note += (" Sweep stream (the two-steps-per-launch kernels' access pattern without "
                 "arithmetic, pylbm_hip_sweep_probe) at D2Q9 4096^2: " + ", ".join(sw)
                 + " GB/s.")
    return "\n".join(rows) + ("\n\n" + note if note else "")


def validation_table(val: dict) -> str:
    """Markdown table of every validation block present in the record."""
    rows = ["| check | reference | result |", "|---|---|---|"]
    for r in val.get("ghia", []):
        e = r["errors"]
        rows.append(f"| cavity Re={r['re']}, {r['n']}², {r['collision']}, {r['backend']} | "
                    f"Ghia et al. 1982 centrelines | L2 = {e['L2']:.4f}, Linf = "
                    f"{e['Linf']:.4f} (units of U), steps {r['steps']}, converged "
                    f"{r['converged']} |")
    sw = val.get("shear_wave", [])
    if sw:
        backends = ", ".join(sorted({r["backend"] for r in sw}))
        rows.append(f"| shear-wave decay, tau {min(r['tau'] for r in sw)}–"
                    f"{max(r['tau'] for r in sw)}, n = {sw[0]['n']} ({backends}) | "
                    f"nu = (tau − 1/2)/3 | max rel. error "
                    f"{100 * max(r['rel_error'] for r in sw):.2f}% |")
    for coll, c in val.get("poiseuille", {}).items():
        if "runs" in c:
            hs = ", ".join(str(r["h"]) for r in c["runs"])
            order = "—" if c.get("order") is None else f"{c['order']:.3f}"
            rows.append(f"| Poiseuille {coll.upper()}, H = {hs} ({c['runs'][0]['backend']}) | "
                        f"analytic parabola | order {order}, max L2 "
                        f"{max(r['L2'] for r in c['runs']):.1e} |")
        else:
            rows.append(f"| Poiseuille {coll}, H = {c['h']} ({c['backend']}) | analytic "
                        f"parabola | L2 {c['L2']:.1e}, Linf {c['Linf']:.1e} |")
    k = val.get("karman")
    if k:
        ref = k.get("reference", {})
        span = ", ".join(f"{name} {lo}–{hi}" for name, (lo, hi) in ref.items())
        rows.append(f"| cylinder 2D-2, D = {k['d']} ({k['backend']}) | Schäfer–Turek 1996"
                    f"{': ' + span if span else ''} | St = {k['St']:.4f}, cD,max = "
                    f"{k['cd_max']:.3f}, cL,max = {k['cl_max']:.3f} |")
    tg = val.get("taylor_green_3d", [])
    for b in sorted({r["backend"] for r in tg}):
        sel = [r for r in tg if r["backend"] == b]
        planes = "/".join(r["plane"] for r in sel)
        rows.append(f"| D3Q19 Taylor–Green, {sel[0]['n']}³, {planes} ({b}) | "
                    f"nu = (tau − 1/2)/3 | max rel. error "
                    f"{100 * max(r['rel_error'] for r in sel):.2f}% |")
    eq = val.get("equivalence", [])
    for st in sorted({r.get("storage", "f32") for r in eq}):
        sel = [r for r in eq if r.get("storage", "f32") == st]
        rows.append(f"| equivalence, {st} storage, {len(sel)} runs | float64 solver | "
                    f"max abs Δf {max(r['max_abs_diff'] for r in sel):.2e} |")
    return "\n".join(rows)


def render(bench: dict | None = None, val: dict | None = None) -> str:
    """The whole report as Markdown."""
    parts = []
    if bench is not None:
        parts += ["## Headline", "", *[f"- {h}" for h in headline(bench)], "",
                  "## Performance", "", performance_table(bench), ""]
    if val is not None:
        parts += ["## Validation", "", validation_table(val), ""]
    return "\n".join(parts)


_BLOCK = re.compile(r"(<!-- pylbm_amd:(?P<name>[\w-]+) -->\n).*?(\n?<!-- /pylbm_amd:(?P=name) -->)",
                    re.S)


def fill_blocks(text: str, blocks: dict[str, str]) -> str:
    """Replace the body of every ``<!-- pylbm_amd:NAME -->`` block named in ``blocks``;
    a name without its block in ``text`` raises KeyError."""
    seen = set()

    def sub(m: re.Match) -> str:
        name = m.group("name")
        if name not in blocks:
            return m.group(0)
        seen.add(name)
        return m.group(1) + blocks[name].strip() + "\n" + m.group(3).lstrip("\n")

    out = _BLOCK.sub(sub, text)
    missing = set(blocks) - seen
    if missing:
        raise KeyError(f"no block for {sorted(missing)}")
    return out


def blocks_for(bench: dict | None, val: dict | None) -> dict[str, str]:
    out = {}
    if bench is not None:
        out["headline"] = "\n".join(f"- {h}" for h in headline(bench))
        out["performance"] = performance_table(bench)
    if val is not None:
        out["validation"] = validation_table(val)
    return out
