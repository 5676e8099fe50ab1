"""LaTeX macros and figures for the paper, from benchmark and validation records.

The paper side of the reference's results pipeline (figures/make_figures.py:89-296 draws
the figures, :359-553 writes ``paper/results.tex``): every number the paper prints is a
``\\newcommand`` computed here from the records, and every figure is drawn from the same
records.  The inputs are the records ``pylbm_amd.report`` reads (reference schema):

- a benchmark record (``runs`` of backend, variant, lattice, collision, size, median, min,
  max, bytes_per_unit, effective_GBs, pct_copy_bw, pct_spec_bw, plus ``machine`` and
  ``ceiling``), as ``python -m pylbm_amd bench --out`` writes it or as converted from a
  ``bench.py`` line;
- a validation record (ghia, shear_wave, poiseuille, karman, taylor_green_3d,
  equivalence), as ``python -m pylbm_amd validate --out`` writes it;
- optionally a head-to-head record (``runs`` of code, variant, lattice, size, ...;
  ``counterparts`` with the pinned commits) in the reference's results/counterparts.json
  layout.

The macro names, their number formats (thousands grouped with ``{,}``, fixed decimals per
macro) and the labels of the reference's rungs are the interface of the reference's
paper/paper.tex and stay as they are, so the reference's own records reproduce its
results.tex macro for macro.  The hip rungs get labels and rung macros of their own.
"""

from __future__ import annotations

import math
from pathlib import Path
from typing import Any, Callable, Iterable

#: (backend, variant) -> label, in ladder order.  The reference's rungs first
#: (figures/make_figures.py:34-52), then the hip ones.
RUNG_LABELS: dict[tuple[str, str], str] = {
    ("numpy", "roll"): "NumPy (fp64)",
    ("torch-eager", "roll"): "torch eager, roll",
    ("torch-eager", "gather"): "torch eager, gather",
    ("torch-compile", "original"): "original code (torch.compile)",
    ("torch-compile", "roll"): "torch.compile, roll",
    ("torch-compile", "gather"): "torch.compile, gather",
    ("numba", "fused"): "Numba-CUDA fused",
    ("warp", "fused"): "Warp fused",
    ("triton", "fused"): "Triton fused",
    ("cupy-cuda-c", "fused"): "CUDA C fused",
    ("triton", "fused-aa"): "Triton fused AA",
    ("cupy-cuda-c", "fused-aa"): "CUDA C fused AA",
    ("triton", "fused-fp16"): "Triton fused fp16",
    ("triton", "fused-bf16"): "Triton fused bf16",
    ("cupy-cuda-c", "fused-fp16"): "CUDA C fused fp16",
    ("triton", "fused-aa-fp16"): "Triton AA fp16",
    ("cupy-cuda-c", "fused-aa-fp16"): "CUDA C AA fp16",
    ("hip", "gather"): "HIP gather",
    ("hip", "fused"): "HIP fused",
    ("hip", "fused-f32s"): "HIP fused f32s",
    ("hip", "fused-aa"): "HIP fused AA",
    ("hip", "fused-fp16"): "HIP fused fp16",
    ("hip", "fused-bf16"): "HIP fused bf16",
    ("hip", "fused-aa-fp16"): "HIP AA fp16",
    ("hip", "fused-aa-bf16"): "HIP AA bf16",
    ("hip", "fused-2step"): "HIP 2 steps/launch",
    ("hip", "fused-2step-fp16"): "HIP 2 steps/launch fp16",
    ("hip", "fused-2step-bf16"): "HIP 2 steps/launch bf16",
}

#: bar colours per backend
COLOURS = {"numpy": "#7f7f7f", "torch-eager": "#bcbd22", "torch-compile": "#ff7f0e",
           "numba": "#8c564b", "warp": "#17becf", "triton": "#1f77b4",
           "cupy-cuda-c": "#2ca02c", "hip": "#d62728"}

#: the fused pull kernel whose largest size sets a ladder's reference size, by preference
PRIMARY = (("hip", "fused"), ("triton", "fused"))

#: rungs with macros of their own: (backend, variant) -> macro stem
RUNG_MACROS = {("torch-compile", "roll"): "compileRoll", ("torch-eager", "roll"): "eagerRoll",
               ("numpy", "roll"): "numpyRef", ("triton", "fused"): "tritonFused",
               ("warp", "fused"): "warpFused", ("numba", "fused"): "numbaFused",
               ("cupy-cuda-c", "fused"): "cudaFused", ("hip", "fused"): "hipFused",
               ("hip", "fused-aa"): "hipAA", ("hip", "fused-fp16"): "hipHalf",
               ("hip", "fused-aa-fp16"): "hipAAHalf"}

#: Ghia runs with macros: solver name -> tag in the macro name
GHIA_TAGS = {"triton/fused": "", "triton/fused-fp16": "Hal", "triton/fused-bf16": "Bf",
             "torch-compile/original": "Orig", "triton/fused-aa": "AA",
             "cupy-cuda-c/fused-aa-fp16": "CudaAAHal", "hip/fused": "",
             "hip/fused-fp16": "Hal", "hip/fused-bf16": "Bf", "hip/fused-aa": "AA",
             "hip/fused-aa-fp16": "AAHal"}
GHIA_RE = {100: "One", 400: "Four", 1000: "Thou"}
#: the grid each Reynolds number is quoted on; other grids carry their size in the name
GHIA_GRID = {100: 128, 400: 256, 1000: 256}
GRID_WORDS = {64: "Sixtyfour", 128: "Onetwentyeight", 256: "Twofiftysix", 512: "Fivetwelve",
              1024: "Tentwentyfour"}

#: head-to-head rungs: (code, variant) -> (macro tag, label, colour, hatch)
COUNTERPARTS = {
    ("pylbm", "triton fused"): ("PyTriton", "pylbm Triton fused", COLOURS["triton"], ""),
    ("pylbm", "cupy-cuda-c fused-aa"): ("PyCudaAA", "pylbm CUDA C fused AA",
                                        COLOURS["cupy-cuda-c"], ""),
    ("FluidX3D", "FP32/FP32"): ("FxThirtyTwo", "FluidX3D FP32/FP32", "#000000", ""),
    ("Lettuce", "native CUDA"): ("LtNative", "Lettuce native CUDA", "#555555", ""),
    ("Lettuce", "PyTorch"): ("LtTorch", "Lettuce PyTorch", "#aaaaaa", ""),
    ("pylbm", "triton fused-fp16"): ("PyTritonHalf", "pylbm Triton fp16", COLOURS["triton"],
                                     "///"),
    ("pylbm", "cupy-cuda-c fused-aa-fp16"): ("PyCudaAAHalf", "pylbm CUDA C AA fp16",
                                             COLOURS["cupy-cuda-c"], "///"),
    ("FluidX3D", "FP32/FP16S"): ("FxSixteenS", "FluidX3D FP32/FP16S", "#000000", "///"),
    ("FluidX3D", "FP32/FP16C"): ("FxSixteenC", "FluidX3D FP32/FP16C", "#000000", "..."),
}
LATTICE_WORD = {"D2Q9": "Two", "D3Q19": "Three"}
#: the head-to-head ratios: macro stem -> (numerator rungs, denominator rung); the
#: numerator is the faster of its rungs
CP_RATIOS = {
    "PyOverFx": ((("pylbm", "triton fused"), ("pylbm", "cupy-cuda-c fused-aa")),
                 ("FluidX3D", "FP32/FP32")),
    "PyHalfOverFxS": ((("pylbm", "triton fused-fp16"), ("pylbm", "cupy-cuda-c fused-aa-fp16")),
                      ("FluidX3D", "FP32/FP16S")),
    "PyOverLtNative": ((("pylbm", "triton fused"), ("pylbm", "cupy-cuda-c fused-aa")),
                       ("Lettuce", "native CUDA")),
    "PyOverLtTorch": ((("pylbm", "triton fused"), ("pylbm", "cupy-cuda-c fused-aa")),
                      ("Lettuce", "PyTorch")),
}

Rung = tuple[str, str]


# ----------------------------------------------------------------------------- numbers
def grouped(x: float, decimals: int = 0) -> str:
    """``x`` with ``decimals`` digits and thousands grouped by ``{,}`` (LaTeX-safe comma)."""
    return f"{x:,.{decimals}f}".replace(",", "{,}")


def _pct0(run: dict) -> str:
    return f"{run['pct_copy_bw']:.0f}"


def nodes(run: dict) -> int:
    return math.prod(int(s) for s in run["size"])


def label(key: Rung) -> str:
    return RUNG_LABELS.get(key, f"{key[0]}/{key[1]}")


def _is16(key: Rung) -> bool:
    return "16" in key[1]


# ----------------------------------------------------------------------------- selection
def _runs(bench: dict, lattice: str, collision: str = "bgk") -> list[dict]:
    return [r for r in bench.get("runs", [])
            if r.get("lattice") == lattice and r.get("collision") == collision
            and not r.get("workload")]


def reference_nodes(runs: list[dict]) -> int | None:
    """Largest lattice measured by the primary fused pull kernel."""
    for key in PRIMARY:
        sizes = [nodes(r) for r in runs if (r["backend"], r["variant"]) == key]
        if sizes:
            return max(sizes)
    return None


def paper_ladder(bench: dict, lattice: str) -> dict[Rung, dict]:
    """The BGK ladder of the paper: per rung its run at the reference size, or its largest
    run where it did not reach that size (slow rungs).  Ties: the faster run."""
    runs = _runs(bench, lattice)
    ref = reference_nodes(runs)
    sel: dict[Rung, dict] = {}
    for r in runs:
        key = (r["backend"], r["variant"])
        cur = sel.get(key)
        at_ref = ref is not None and nodes(r) == ref
        if cur is None:
            sel[key] = r
            continue
        cur_at_ref = ref is not None and nodes(cur) == ref
        if (at_ref, nodes(r), r["median"]) > (cur_at_ref, nodes(cur), cur["median"]):
            sel[key] = r
    order = [k for k in RUNG_LABELS if k in sel] + sorted(k for k in sel if k not in RUNG_LABELS)
    return {k: sel[k] for k in order}


def paper_ladders(bench: dict) -> dict[str, dict[Rung, dict]]:
    return {lat: sel for lat in ("D2Q9", "D3Q19") if (sel := paper_ladder(bench, lat))}


def _best(sel: dict[Rung, dict], keep: Callable[[Rung], bool]) -> tuple[Rung, dict] | None:
    cand = [(k, r) for k, r in sel.items() if keep(k)]
    return max(cand, key=lambda kr: kr[1]["median"]) if cand else None


# ------------------------------------------------------------------------------- macros
def bench_macros(bench: dict, source: str) -> dict[str, str]:
    """Ceilings, the ladder table and the headline rates (reference
    figures/make_figures.py:400-475)."""
    ceil = bench["ceiling"]
    copy_bw, spec_bw = ceil["copy_bandwidth_GBs"], ceil["spec_bandwidth_GBs"]
    machine = bench.get("machine", {})
    m = {"benchSource": source.replace("_", r"\_"), "copyBW": grouped(copy_bw),
         "specBW": grouped(spec_bw), "gpuName": machine.get("gpu", "GPU"),
         "benchDate": machine.get("date", "")[:10]}
    # MLUPS at copy bandwidth: D2Q9 fp32 (72 B/update), 16-bit (36 B), D3Q19 fp32 (152 B)
    for name, bpu in (("ceilFpThirtyTwo", 72), ("ceilSixteen", 36), ("ceilThreeD", 152)):
        m[name] = grouped(copy_bw * 1e3 / bpu)
    ladders = paper_ladders(bench)
    blocks = []
    for lattice, sel in ladders.items():
        rows = []
        for key, r in sel.items():
            size = "$" + r"\times".join(str(s) for s in r["size"]) + "$"
            rows.append(f"{lattice} & {label(key)} & {size} & {grouped(r['median'])} & "
                        f"{r['bytes_per_unit']} & {grouped(r['effective_GBs'])} & "
                        f"{r['pct_copy_bw']:.0f} & {r['pct_spec_bw']:.0f} \\\\")
        blocks.append("\n".join(rows))
    m["ladderRows"] = "\n\\midrule\n".join(blocks)

    d2 = ladders.get("D2Q9", {})
    b32 = _best(d2, lambda k: not _is16(k) and k[0] != "numpy")
    b16 = _best(d2, _is16)
    ball = _best(d2, lambda k: True)
    if b32:
        k, r = b32
        m.update(bestFpThirtyTwoMLUPS=grouped(r["median"]), bestFpThirtyTwoName=label(k),
                 bestFpThirtyTwoPct=_pct0(r), bestFpThirtyTwoGBs=grouped(r["effective_GBs"]))
    if b16:
        k, r = b16
        m.update(bestSixteenMLUPS=grouped(r["median"]), bestSixteenName=label(k),
                 bestSixteenPct=_pct0(r))
    if ball:
        k, r = ball
        m.update(bestMLUPS=grouped(r["median"]), bestName=label(k), bestPctCopy=_pct0(r),
                 bestPctSpec=f"{r['pct_spec_bw']:.0f}", bestGBs=grouped(r["effective_GBs"]))
    orig = d2.get(("torch-compile", "original"))
    if orig and b32:
        m.update(origMLUPS=grouped(orig["median"]), origPct=f"{orig['pct_copy_bw']:.1f}",
                 speedupFpThirtyTwo=f"{b32[1]['median'] / orig['median']:.1f}",
                 speedupBest=f"{ball[1]['median'] / orig['median']:.1f}")
    for key, stem in RUNG_MACROS.items():
        if key in d2:
            m[stem + "MLUPS"] = grouped(d2[key]["median"])
            m[stem + "Pct"] = _pct0(d2[key])

    d3 = ladders.get("D3Q19", {})
    b3 = _best(d3, lambda k: not _is16(k) and k[0] != "numpy")
    if b3:
        k, r = b3
        m.update(bestThreeDMLUPS=grouped(r["median"]), bestThreeDName=label(k),
                 bestThreeDPct=_pct0(r))
    if ("torch-compile", "roll") in d3:
        m["compileRollThreeDPct"] = _pct0(d3[("torch-compile", "roll")])
    b3h = _best(d3, _is16)
    if b3h:
        m.update(bestThreeDSixteenMLUPS=grouped(b3h[1]["median"]),
                 bestThreeDSixteenPct=_pct0(b3h[1]))

    # the primary fused kernel under the other collision operators (D2Q9, largest lattice)
    for backend, variant in PRIMARY:
        other: dict[str, dict] = {}
        for r in bench.get("runs", []):
            if ((r["backend"], r["variant"]) != (backend, variant) or r["collision"] == "bgk"
                    or r.get("lattice") != "D2Q9" or r.get("workload")):
                continue
            cur = other.get(r["collision"])
            if cur is None or (nodes(r), r["median"]) > (nodes(cur), cur["median"]):
                other[r["collision"]] = r
        for coll, r in other.items():
            m[f"{backend}{coll.upper()}Pct"] = _pct0(r)
            m[f"{backend}{coll.upper()}MLUPS"] = grouped(r["median"])
    return m


def validation_macros(val: dict) -> dict[str, str]:
    """Ghia errors, viscosity, Poiseuille order, cylinder, Taylor-Green and equivalence
    numbers (reference figures/make_figures.py:476-538)."""
    m: dict[str, str] = {}
    for r in val.get("ghia", []):
        tag = GHIA_TAGS.get(r["backend"])
        if tag is None or r["re"] not in GHIA_RE:
            continue
        coll = "" if r["collision"] == "bgk" else r["collision"].upper()
        grid = "" if GHIA_GRID[r["re"]] == r["n"] else GRID_WORDS.get(r["n"], "")
        name = f"ghia{GHIA_RE[r['re']]}{coll}{tag}{grid}"
        m[name + "Ltwo"] = f"{r['errors']['L2']:.4f}"
        m[name + "Linf"] = f"{r['errors']['Linf']:.4f}"
        m[name + "N"] = str(r["n"])
    sw = val.get("shear_wave", [])
    if sw:
        taus = [r["tau"] for r in sw]
        mid = [r["rel_error"] for r in sw if 0.6 <= r["tau"] <= 1.6]
        m.update(shearMaxErr=f"{100 * max(r['rel_error'] for r in sw):.2f}",
                 shearN=str(sw[0]["n"]), shearTauMin=f"{min(taus):.2f}",
                 shearTauMax=f"{max(taus):.2f}",
                 shearMidErr=f"{100 * max(mid):.2f}" if mid else "--")
    po = val.get("poiseuille", {})
    bgk_runs = po.get("bgk", {}).get("runs", [])
    if bgk_runs:
        if po["bgk"].get("order") is not None:
            m["poisOrder"] = f"{po['bgk']['order']:.3f}"
        m["poisBgkFinest"] = f"{max(bgk_runs, key=lambda r: r['h'])['L2']:.1e}"
    if po.get("trt", {}).get("runs"):
        m["poisTrtMax"] = f"{max(r['L2'] for r in po['trt']['runs']):.0e}"
    if "bgk_fp32_h32" in po:
        m["poisFpThirtyTwo"] = f"{po['bgk_fp32_h32']['L2']:.1e}"
        h32 = [r for r in bgk_runs if r["h"] == 32]
        if h32:
            m["poisFpSixtyFourHThirtyTwo"] = f"{h32[0]['L2']:.1e}"
    k = val.get("karman")
    if k:
        m.update(karmanSt=f"{k['St']:.4f}", karmanCd=f"{k['cd_max']:.3f}",
                 karmanCl=f"{k['cl_max']:.3f}", karmanD=str(k["d"]),
                 karmanTau=f"{k['tau']:.3f}", karmanNx=str(k["shape"][0]),
                 karmanNy=str(k["shape"][1]))
    tg = val.get("taylor_green_3d", [])
    if tg:
        dev = [r for r in tg if not r["backend"].startswith("numpy")] or tg
        m.update(tgvMaxErr=f"{100 * max(r['rel_error'] for r in dev):.2f}",
                 tgvN=str(dev[0]["n"]))
    eq = val.get("equivalence", [])
    if eq:
        for name, storage in (("eqMaxFpThirtyTwo", "f32"), ("eqMaxFpSixteen", "f16"),
                              ("eqMaxBf", "bf16")):
            d = [r["max_abs_diff"] for r in eq if r.get("storage", "f32") == storage]
            m[name] = f"{max(d):.1e}" if d else "--"
        m["eqCount"] = str(len({(r["backend"], r["variant"]) for r in eq}))
    return m


def counterpart_rungs(cp: dict) -> dict[str, dict[Rung, dict]]:
    """Per lattice and known head-to-head rung: its run on the largest lattice."""
    out: dict[str, dict[Rung, dict]] = {}
    for r in cp.get("runs", []):
        key = (r["code"], r["variant"])
        if key not in COUNTERPARTS:
            continue
        sel = out.setdefault(r["lattice"], {})
        if key not in sel or nodes(r) > nodes(sel[key]):
            sel[key] = r
    return {lat: {k: sel[k] for k in COUNTERPARTS if k in sel} for lat, sel in out.items()}


def _size_tex(size: list[int]) -> str:
    if len(set(size)) == 1:
        return f"${size[0]}^{len(size)}$"
    return "$" + r"\times".join(str(s) for s in size) + "$"


def counterpart_macros(cp: dict) -> dict[str, str]:
    """Head-to-head rates, ratios and pinned commits (reference
    figures/make_figures.py:359-393)."""
    m: dict[str, str] = {}
    for lat, sel in counterpart_rungs(cp).items():
        word = LATTICE_WORD[lat]
        for key, r in sel.items():
            tag = COUNTERPARTS[key][0]
            m[f"cp{tag}{word}MLUPS"] = grouped(r["median"])
            m[f"cp{tag}{word}Pct"] = _pct0(r)
            m[f"cp{tag}{word}Size"] = _size_tex(r["size"])
        for stem, (nums, den) in CP_RATIOS.items():
            num = max((sel[k]["median"] for k in nums if k in sel), default=None)
            if num and den in sel and sel[den]["median"]:
                ratio = num / sel[den]["median"]
                m[f"cp{stem}{word}"] = f"{ratio:.2f}" if ratio < 10 else f"{ratio:.0f}"
    codes = cp.get("counterparts", {})
    fx, lt = codes.get("FluidX3D", {}), codes.get("Lettuce", {})
    if fx:
        m["cpFxCommit"] = (fx.get("commit") or fx["pinned"])[:7]
    if lt:
        m["cpLtCommit"] = (lt.get("commit") or lt["pinned"])[:7]
        m["cpLtVersion"] = str(lt.get("lettuce", "--")).split("+")[0]
        m["cpLtTorch"] = str(lt.get("torch", "--"))
    m["cpDate"] = cp.get("machine", {}).get("date", "")[:10]
    m["cpPlan"] = cp.get("plan", "--")
    return m


def macros(bench: dict | None, val: dict | None, cp: dict | None = None,
           source: str = "bench.json") -> dict[str, str]:
    m: dict[str, str] = {}
    if bench is not None:
        m.update(bench_macros(bench, source))
    if val is not None:
        m.update(validation_macros(val))
    if cp is not None:
        m.update(counterpart_macros(cp))
    return m


def results_tex(bench: dict | None, val: dict | None, cp: dict | None = None,
                source: str = "bench.json") -> str:
    """``results.tex``: the provisional flags and one ``\\newcommand`` per macro, sorted."""
    prov = bool(bench.get("provisional", True)) if bench is not None else True
    cp_prov = bool(cp.get("provisional", True)) if cp is not None else True
    lines = ["% Generated by `python -m pylbm_amd report --tex`; edits are overwritten.",
             f"% benchmark: {source}" + ("; validation record" if val is not None else "")
             + ("; head-to-head record" if cp is not None else ""),
             r"\newif\ifprovisional", r"\provisionaltrue" if prov else r"\provisionalfalse",
             r"\newif\ifcpprovisional",
             r"\cpprovisionaltrue" if cp_prov else r"\cpprovisionalfalse"]
    for name, value in sorted(macros(bench, val, cp, source).items()):
        lines.append(f"\\newcommand{{\\{name}}}{{{value}}}")
    return "\n".join(lines) + "\n"


_NEWCOMMAND = "\\newcommand{\\"


def parse_results_tex(text: str) -> dict[str, Any]:
    """Macros and flags of a results.tex file: ``{"macros": {name: value}, "provisional":
    bool, "cpprovisional": bool}``.  Values may span lines (the ladder rows)."""
    out: dict[str, Any] = {"macros": {}}
    i = 0
    while (j := text.find(_NEWCOMMAND, i)) >= 0:
        k = text.index("}", j)
        name = text[j + len(_NEWCOMMAND):k]
        assert text[k + 1] == "{", name
        depth, p = 1, k + 2
        while depth:
            depth += {"{": 1, "}": -1}.get(text[p], 0)
            p += 1
        out["macros"][name] = text[k + 2:p - 1]
        i = p
    for flag in ("provisional", "cpprovisional"):
        if f"\\{flag}true" in text:
            out[flag] = True
        elif f"\\{flag}false" in text:
            out[flag] = False
    return out


# ------------------------------------------------------------------------------ figures
def _plt():
    import matplotlib

    matplotlib.use("Agg")
    import matplotlib.pyplot as plt

    plt.rcParams.update({"font.family": "serif", "font.size": 9, "axes.spines.top": False,
                         "axes.spines.right": False, "figure.dpi": 150,
                         "savefig.bbox": "tight"})
    return plt


WIDTH = 7.0  # inches: a full text width


def _save(fig, out_dir: Path, name: str, formats: Iterable[str]) -> list[Path]:
    import matplotlib.pyplot as plt

    paths = []
    for ext in formats:
        p = out_dir / f"{name}.{ext}"
        fig.savefig(p)
        paths.append(p)
    plt.close(fig)
    return paths


#: hip rungs share a backend, so their figures colour by variant
HIP_COLOURS = {"gather": "#8c564b", "fused": "#d62728", "fused-f32s": "#ff9896",
               "fused-aa": "#ff7f0e", "fused-fp16": "#1f77b4", "fused-bf16": "#9467bd",
               "fused-aa-fp16": "#2ca02c", "fused-aa-bf16": "#17becf", "fused-2step": "#7f7f7f",
               "fused-2step-fp16": "#bcbd22", "fused-2step-bf16": "#e377c2"}


def _colour(key: Rung) -> str:
    if key[0] == "hip":
        return HIP_COLOURS.get(key[1], COLOURS["hip"])
    return COLOURS.get(key[0], "#444444")


def fig_ladder(bench: dict, out_dir: Path, formats: Iterable[str] = ("pdf", "png")):
    """MLUPS per rung (log scale) at the reference size with the fp32 and 16-bit copy
    ceilings, one panel per lattice (reference figures/make_figures.py:89-137)."""
    plt = _plt()
    bw = bench["ceiling"]["copy_bandwidth_GBs"]
    ladders = paper_ladders(bench)
    lats = list(ladders)
    if not lats:
        return []
    fig, axes = plt.subplots(1, len(lats), figsize=(WIDTH, 3.6), squeeze=False,
                             gridspec_kw={"width_ratios": [1.5, 1][:len(lats)]})
    for ax, lat in zip(axes[0], lats):
        sel, q = ladders[lat], 9 if lat == "D2Q9" else 19
        keys = list(sel)
        med = [sel[k]["median"] for k in keys]
        y = list(range(len(keys)))[::-1]
        err = [[sel[k]["median"] - sel[k]["min"] for k in keys],
               [sel[k]["max"] - sel[k]["median"] for k in keys]]
        ax.barh(y, med, color=[_colour(k) for k in keys], xerr=err,
                error_kw={"lw": 0.6, "capsize": 1.5})
        ax.set_yticks(y, [f"{label(k)} ({sel[k]['size'][0]}$^{len(sel[k]['size'])}$)"
                          for k in keys], fontsize=7)
        ax.set_xscale("log")
        for bpu, ls, txt in ((8 * q, "--", "fp32"), (4 * q, ":", "16-bit")):
            c = bw * 1e3 / bpu
            ax.axvline(c, color="k", ls=ls, lw=0.9, zorder=0)
            ax.text(c, 1.0, txt, transform=ax.get_xaxis_transform(), fontsize=6.5,
                    ha="right" if txt == "fp32" else "left", va="bottom")
        for yi, v in zip(y, med):
            ax.text(v * 1.12, yi, f"{v:,.0f}", va="center", fontsize=6,
                    bbox={"boxstyle": "square,pad=0.1", "fc": "white", "ec": "none",
                          "alpha": 0.85})
        ref = reference_nodes(_runs(bench, lat)) or max(nodes(r) for r in sel.values())
        dim = 2 if q == 9 else 3
        ax.set_title(f"{lat} cavity (fused kernels at {round(ref ** (1 / dim))}$^{dim}$)",
                     fontsize=8, pad=10)
        ax.set_xlabel("MLUPS (log scale)")
        ax.set_xlim(left=max(1.0, 0.5 * min(med)), right=4 * bw * 1e3 / (4 * q))
    fig.tight_layout()
    return _save(fig, out_dir, "ladder", formats)


def fig_scaling(bench: dict, out_dir: Path, formats: Iterable[str] = ("pdf", "png")):
    """MLUPS against lattice nodes for every rung measured at two sizes or more, with the
    copy ceilings (reference figures/make_figures.py:140-174)."""
    plt = _plt()
    bw = bench["ceiling"]["copy_bandwidth_GBs"]
    fig, axes = plt.subplots(1, 2, figsize=(WIDTH, 2.8))
    for ax, lat, q in ((axes[0], "D2Q9", 9), (axes[1], "D3Q19", 19)):
        groups: dict[Rung, list[dict]] = {}
        for r in _runs(bench, lat):
            groups.setdefault((r["backend"], r["variant"]), []).append(r)
        for key, rs in groups.items():
            if len(rs) < 2 or key[0] == "numpy":
                continue
            rs = sorted(rs, key=nodes)
            v = key[1]
            style = "-." if "bf16" in v else ":" if "fp16" in v else "--" if "aa" in v else "-"
            marker = ("v" if "aa" in v else "d") if _is16(key) else "x" if "aa" in v else \
                {"gather": "s", "original": "^"}.get(v, "o")
            ax.plot([nodes(r) for r in rs], [r["median"] for r in rs], style, marker=marker,
                    ms=3, lw=1, color=_colour(key), label=label(key))
        for bpu, ls in ((8 * q, "--"), (4 * q, ":")):
            ax.axhline(bw * 1e3 / bpu, color="k", ls=ls, lw=0.8)
        ax.set_xscale("log")
        ax.set_yscale("log")
        ax.set_xlabel("lattice nodes")
        ax.set_ylabel("MLUPS")
        ax.set_title(lat, fontsize=8)
    h, lab = axes[0].get_legend_handles_labels()
    fig.legend(h, lab, fontsize=6, loc="center left", bbox_to_anchor=(1.0, 0.5),
               frameon=False)
    fig.tight_layout()
    return _save(fig, out_dir, "scaling", formats)


def fig_ghia(val: dict, out_dir: Path, formats: Iterable[str] = ("pdf", "png")):
    """Cavity centrelines of the primary BGK runs (finest grid per Re) over the Ghia et al.
    (1982) table (reference figures/make_figures.py:177-205)."""
    from .validation.ghia import reference

    primary = {"triton/fused", "numpy/roll", "hip/fused"}
    by_re: dict[int, dict] = {}
    for r in val.get("ghia", []):
        if r["collision"] != "bgk" or r["backend"] not in primary or r["re"] not in GHIA_RE:
            continue
        if r["re"] not in by_re or r["n"] > by_re[r["re"]]["n"]:
            by_re[r["re"]] = r
    if not by_re:
        return []
    plt = _plt()
    fig, axes = plt.subplots(1, 2, figsize=(WIDTH, 3.0))
    colours = {100: "#1f77b4", 400: "#2ca02c", 1000: "#d62728"}
    for re in sorted(by_re):
        r, ref, c = by_re[re], reference(re), colours[re]
        p = r["profile"]
        axes[0].plot(p["u"], p["y"], color=c, lw=1, label=f"Re={re} ({r['n']}$^2$)")
        axes[0].plot(ref["u"], ref["y"], "o", mfc="none", color=c, ms=3.5)
        axes[1].plot(p["x"], p["v"], color=c, lw=1)
        axes[1].plot(ref["x"], ref["v"], "o", mfc="none", color=c, ms=3.5)
    axes[0].plot([], [], "o", mfc="none", color="k", ms=3.5, label="Ghia et al. (1982)")
    axes[0].set_xlabel("$u/U$ on $x=1/2$")
    axes[0].set_ylabel("$y$")
    axes[1].set_xlabel("$x$")
    axes[1].set_ylabel("$v/U$ on $y=1/2$")
    axes[0].legend(fontsize=7, frameon=False)
    fig.tight_layout()
    return _save(fig, out_dir, "ghia", formats)


def fig_viscosity(val: dict, out_dir: Path, formats: Iterable[str] = ("pdf", "png")):
    """Measured against theoretical viscosity over tau, and the relative error
    (reference figures/make_figures.py:208-226)."""
    sw = sorted(val.get("shear_wave", []), key=lambda r: r["tau"])
    if not sw:
        return []
    plt = _plt()
    fig, axes = plt.subplots(1, 2, figsize=(WIDTH, 2.6))
    tau = [r["tau"] for r in sw]
    c = COLOURS["hip"] if sw[0]["backend"].startswith("hip") else COLOURS["triton"]
    axes[0].plot(tau, [r["nu_theory"] for r in sw], "k--", lw=0.8, label=r"theory $(\tau-1/2)/3$")
    axes[0].plot(tau, [r["nu_measured"] for r in sw], "o", color=c, ms=3.5,
                 label=f"measured ({sw[0]['backend']}, $n={sw[0]['n']}$)")
    axes[0].set_xlabel(r"$\tau$")
    axes[0].set_ylabel(r"$\nu$ (lattice units)")
    axes[0].legend(fontsize=7, frameon=False)
    axes[1].semilogy(tau, [max(r["rel_error"], 1e-8) for r in sw], "o-", ms=3.5, color=c, lw=0.8)
    axes[1].set_xlabel(r"$\tau$")
    axes[1].set_ylabel(r"$|\nu_{\rm meas}-\nu|/\nu$")
    fig.tight_layout()
    return _save(fig, out_dir, "viscosity", formats)


def fig_poiseuille(val: dict, out_dir: Path, formats: Iterable[str] = ("pdf", "png")):
    """Grid convergence of the channel error with a slope -2 guide, and one profile
    against the parabola (reference figures/make_figures.py:229-258)."""
    po = val.get("poiseuille", {})
    if not po.get("bgk", {}).get("runs"):
        return []
    plt = _plt()
    fig, axes = plt.subplots(1, 2, figsize=(WIDTH, 2.6))
    for coll, marker, c in (("bgk", "o", COLOURS["numpy"]), ("trt", "s", COLOURS["cupy-cuda-c"])):
        rs = sorted(po.get(coll, {}).get("runs", []), key=lambda r: r["h"])
        if not rs:
            continue
        order = po[coll].get("order")
        lab = coll.upper() + ((f" (order {order:.2f})" if order is not None else "")
                              if coll == "bgk" else ", magic $\\Lambda=3/16$")
        axes[0].loglog([r["h"] for r in rs], [max(r["L2"], 1e-16) for r in rs],
                       marker + "-", ms=3.5, lw=0.8, color=c, label=lab)
    rs = sorted(po["bgk"]["runs"], key=lambda r: r["h"])
    h0, e0 = rs[0]["h"], rs[0]["L2"]
    hs = [h0, 8 * h0]
    axes[0].loglog(hs, [e0 * (h / h0) ** -2 for h in hs], "k:", lw=0.8, label="slope $-2$")
    axes[0].set_xlabel("channel width $H$ (nodes)")
    axes[0].set_ylabel("relative $L_2$ error")
    axes[0].legend(fontsize=7, frameon=False)
    r = rs[min(1, len(rs) - 1)]
    ys = [i / 199 for i in range(200)]
    axes[1].plot([4 * y * (1 - y) for y in ys], ys, "k-", lw=0.8, label="analytic")
    axes[1].plot(r["u"], r["y"], "o", ms=3, color=COLOURS["numpy"], label=f"BGK, $H={r['h']}$")
    axes[1].set_xlabel("$u/u_{\\max}$")
    axes[1].set_ylabel("$y/H$")
    axes[1].legend(fontsize=7, frameon=False)
    fig.tight_layout()
    return _save(fig, out_dir, "poiseuille", formats)


def fig_karman(val: dict, out_dir: Path, vorticity: dict | None = None,
               formats: Iterable[str] = ("pdf", "png")):
    """Vorticity of the cylinder wake (solid nodes blank) over the lift and drag histories
    (reference figures/make_figures.py:261-280).  ``vorticity`` holds the arrays
    ``vorticity`` and ``solid`` (``validate --vorticity-out``); without it only the force
    histories are drawn."""
    import numpy as np

    k = val.get("karman")
    if not k:
        return []
    plt = _plt()
    fig = plt.figure(figsize=(WIDTH, 3.2))
    if vorticity is not None:
        w = np.asarray(vorticity["vorticity"], dtype=float)
        w[np.asarray(vorticity["solid"], dtype=bool)] = np.nan
        ax = fig.add_axes([0.0, 0.42, 1.0, 0.58])
        ax.imshow(w, origin="lower", cmap="RdBu_r", vmin=-3, vmax=3, interpolation="bilinear")
        ax.set_axis_off()
    ax2 = fig.add_axes([0.08, 0.08, 0.86, 0.28])
    t = np.asarray(k["t_conv"], dtype=float)
    ax2.plot(t, k["cl"], color=COLOURS["torch-compile"], lw=0.8, label="$c_L$")
    ax2.plot(t, np.asarray(k["cd"], dtype=float) - 3.0, color="k", lw=0.8, label="$c_D-3$")
    ax2.set_xlabel("$t\\,U/D$")
    ax2.legend(fontsize=7, frameon=False, ncol=2, loc="upper right")
    return _save(fig, out_dir, "karman", formats)


def fig_counterparts(cp: dict, out_dir: Path, formats: Iterable[str] = ("pdf", "png")):
    """Head-to-head: % of copy bandwidth per code, MLUPS as labels (reference
    figures/make_figures.py:323-356)."""
    sel = counterpart_rungs(cp)
    lats = [lat for lat in ("D2Q9", "D3Q19") if lat in sel]
    if not lats:
        return []
    plt = _plt()
    fig, axes = plt.subplots(1, len(lats), figsize=(WIDTH, 2.7), sharex=True, squeeze=False)
    for ax, lat in zip(axes[0], lats):
        keys = list(sel[lat])
        ref = max(nodes(sel[lat][k]) for k in keys)
        y = list(range(len(keys)))[::-1]
        for yi, key in zip(y, keys):
            r = sel[lat][key]
            _, lab, c, hatch = COUNTERPARTS[key]
            per = r["pct_copy_bw"] / r["median"]
            ax.barh(yi, r["pct_copy_bw"], color="white" if hatch else c, edgecolor=c,
                    hatch=hatch, lw=0.8,
                    xerr=[[(r["median"] - r["min"]) * per], [(r["max"] - r["median"]) * per]],
                    error_kw={"lw": 0.6, "capsize": 1.5})
            txt = f"{r['median']:,.0f}"
            if nodes(r) != ref:
                txt += f" ({'x'.join(str(s) for s in r['size'])})"
            ax.text(r["max"] * per + 3.5, yi, txt, va="center", fontsize=6)
        ax.set_yticks(y, [COUNTERPARTS[k][1] for k in keys], fontsize=7)
        ax.axvline(100, color="k", ls="--", lw=0.9)
        dim = 2 if lat == "D2Q9" else 3
        ax.set_title(f"{lat}, {round(ref ** (1 / dim))}$^{dim}$ (labels: MLUPS)", fontsize=8)
        ax.set_xlim(0, 135)
    if cp.get("provisional", True):
        fig.text(0.5, 0.5, "PROVISIONAL", color="red", alpha=0.35, fontsize=16,
                 ha="center", va="center", rotation=12)
    fig.supxlabel("% of measured copy bandwidth (minimal traffic model); dashed: roofline",
                  fontsize=8)
    fig.tight_layout()
    return _save(fig, out_dir, "counterparts", formats)


def figures(out_dir: str | Path, bench: dict | None = None, val: dict | None = None,
            cp: dict | None = None, vorticity: dict | None = None,
            formats: Iterable[str] = ("pdf", "png")) -> list[Path]:
    """Every figure the records allow, written to ``out_dir``; returns the files."""
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    formats = tuple(formats)
    paths: list[Path] = []
    if bench is not None:
        paths += fig_ladder(bench, out, formats) + fig_scaling(bench, out, formats)
    if val is not None:
        paths += (fig_ghia(val, out, formats) + fig_viscosity(val, out, formats)
                  + fig_poiseuille(val, out, formats) + fig_karman(val, out, vorticity, formats))
    if cp is not None:
        paths += fig_counterparts(cp, out, formats)
    return paths
