"""The validation suite on the hip backend, written as one JSON record.

The cases and the record layout are those of the reference's suite
(pylbm/validation/runner.py:49-175, results/validation.json): Ghia cavity centrelines,
shear-wave viscosity over a range of tau, Poiseuille grid convergence for BGK and TRT,
the Schaefer-Turek 2D-2 cylinder, and the D3Q19 Taylor-Green decay in three planes.  The
keys and per-case fields match, so a report reads either file.  Every run uses a hip
variant; the reference's float64 NumPy runs (Poiseuille convergence, quick mode) become
hip runs here, the grid-convergence runs on fp32 storage of f - w (``fused-f32s``) at
h = 8, 16, 32, where the fp32 rounding floor stays well below the discretisation error.  The reference's
"equivalence" block (every backend against its float64 solver) has no counterpart in the
package, which ships no float64 solver: that comparison is a test (tests/test_hip_parity.py).
"""

from __future__ import annotations

import dataclasses
import datetime as _dt
import json
import time
from pathlib import Path
from typing import Any, Callable

import numpy as np

CASES = ("ghia", "shear", "poiseuille", "karman", "tgv")

#: (Re, n, collision, variant) of the full Ghia block: the reference's cavity runs, with
#: its fastest-rung variants mapped onto the hip ones
GHIA_FULL = [(100, 128, "bgk", "fused"), (400, 256, "bgk", "fused"),
             (1000, 256, "bgk", "fused"), (1000, 256, "trt", "fused"),
             (1000, 256, "mrt", "fused"), (100, 64, "bgk", "fused"),
             (100, 256, "bgk", "fused"), (100, 128, "bgk", "fused-fp16"),
             (100, 128, "bgk", "fused-bf16"), (1000, 256, "bgk", "fused-aa"),
             (1000, 256, "bgk", "fused-fp16"), (1000, 256, "bgk", "fused-bf16"),
             (1000, 256, "bgk", "fused-aa-fp16")]
GHIA_QUICK = [(100, 64, "bgk", "fused")]
SHEAR_TAUS_FULL = [0.52, 0.55, 0.6, 0.7, 0.8, 0.9, 1.0, 1.2, 1.4, 1.6, 1.8, 1.95]
SHEAR_TAUS_QUICK = [0.55, 0.8, 1.0, 1.5]


def run_suite(quick: bool = False, cases: list[str] | None = None, variant: str | None = None,
              out: str | Path | None = None, vorticity_out: str | Path | None = None,
              log: Callable[[str], Any] = print) -> dict:
    """Run the selected cases (all by default) and return the record; write it to ``out``.

    ``variant`` replaces every case's variant (for instance "fused-fp16" runs the whole
    suite on 16-bit storage); the Karman case needs a two-buffer variant and runs on
    ``gather`` when ``variant`` is None or "fused".  ``vorticity_out`` receives the final
    Karman vorticity (npz, float16) as the reference's suite saves it.
    """
    from . import ghia, karman, poiseuille, shear_wave, taylor_green

    want = set(cases or CASES)
    unknown = want - set(CASES)
    if unknown:
        raise ValueError(f"unknown validation cases {sorted(unknown)}; choose from {CASES}")
    res: dict[str, Any] = {
        "date": _dt.datetime.now().astimezone().isoformat(timespec="seconds"),
        "mode": "quick" if quick else "full", "gpu": True}
    t0 = time.perf_counter()

    if "ghia" in want:
        log("[validate] lid-driven cavity vs Ghia et al. (1982)")
        res["ghia"] = []
        for re, n, coll, v in (GHIA_QUICK if quick else GHIA_FULL):
            r = ghia.run_cavity(re, n, collision=coll, backend="hip", variant=variant or v)
            res["ghia"].append(dataclasses.asdict(r))
            log(f"  Re={re:5d} n={n:4d} {coll} {r.backend:18s} steps={r.steps:7d} "
                f"conv={r.converged} L2={r.errors['L2']:.4f} Linf={r.errors['Linf']:.4f}")

    if "shear" in want:
        log("[validate] shear-wave decay viscosity")
        res["shear_wave"] = []
        n = 32 if quick else 128
        for tau in (SHEAR_TAUS_QUICK if quick else SHEAR_TAUS_FULL):
            r = shear_wave.run_shear_wave(tau, n=n, backend="hip", variant=variant or "fused")
            r.pop("t"), r.pop("amplitude")
            res["shear_wave"].append(r)
            log(f"  tau={tau:.2f} nu={r['nu_theory']:.5f} measured={r['nu_measured']:.5f} "
                f"rel.err={r['rel_error']:.2e} ({r['backend']})")

    if "poiseuille" in want:
        log("[validate] Poiseuille channel")
        # The reference converges these in float64 up to h = 64.  In fp32 arithmetic the
        # steady profile carries a rounding offset that grows with the relaxation time
        # h^2 / nu: with fp32 storage of f - w it is ~1e-5 of u_max up to h = 32 (TRT)
        # against 7e-4 for the BGK discretisation error there, and 1e-4 at h = 64, where it
        # would spoil the fitted order.  Plain fp32 storage stalls near 1e-3 at h = 32.
        hs = (8, 16, 32)
        res["poiseuille"] = {}
        for coll in ("bgk", "trt"):
            c = poiseuille.convergence(hs, collision=coll, backend="hip",
                                       variant=variant or "fused-f32s")
            res["poiseuille"][coll] = c
            errs = ", ".join(f"{r['L2']:.2e}" for r in c["runs"])
            log(f"  {coll}: L2 = [{errs}] order={c['order']}")
        if not quick:
            r = poiseuille.run_poiseuille(32, collision="bgk", backend="hip",
                                          variant=variant or "fused")
            res["poiseuille"]["bgk_fp32_h32"] = r
            log(f"  h=32 ({r['backend']}): L2 = {r['L2']:.2e}")

    if "karman" in want:
        log("[validate] Karman street, Schaefer-Turek 2D-2")
        d = 20 if quick else 40
        v = "gather" if variant in (None, "fused") else variant
        if v.startswith("fused-aa"):
            log(f"  skipped: {v} takes box geometries only")
        else:
            r = karman.run_karman(d=d, backend="hip", variant=v)
            vort, solid = r.pop("_vorticity"), r.pop("_solid")
            if vorticity_out is not None:
                np.savez_compressed(vorticity_out,
                                    vorticity=np.nan_to_num(vort).astype(np.float16),
                                    solid=solid, d=d)
            res["karman"] = r
            log(f"  D={d}: St={r['St']:.4f} cd_max={r['cd_max']:.3f} "
                f"cl_max={r['cl_max']:.3f}")

    if "tgv" in want:
        log("[validate] D3Q19 Taylor-Green decay")
        res["taylor_green_3d"] = []
        runs = [(variant or "fused", 16)] if quick else \
            [(variant or "fused", 64), (variant or "fused-fp16", 64)]
        for v, n in runs:
            for plane in ("xy", "yz", "xz"):
                r = taylor_green.run_taylor_green(n=n, plane=plane, backend="hip", variant=v)
                res["taylor_green_3d"].append(r)
                log(f"  {plane} n={n} {r['backend']}: rel.err nu = {r['rel_error']:.2e}")

    res["wall_time_s"] = time.perf_counter() - t0
    if out is not None:
        Path(out).parent.mkdir(parents=True, exist_ok=True)
        Path(out).write_text(json.dumps(res, indent=1))
        log(f"[validate] wrote {out} in {res['wall_time_s']:.0f} s")
    return res
