"""Command line of the hip backend: ``python -m pylbm_amd <command> ...``.

The command and flag names are the reference's (pylbm/cli.py:98-144), so a script written
for the ``pylbm`` console script runs the gfx950 kernels by changing the program name:

* ``run cavity|cavity3d|karman --backend hip --variant V`` -- one case, measured on the
  device;
* ``validate``, ``bench`` -- the validation suite and the benchmark plans of this package;
* ``backends`` -- every registered entry, whether it runs here, its storage and layout;
* ``report`` -- Markdown tables and README blocks from ``bench`` / ``validate`` records;
  with ``--tex`` the paper's ``results.tex`` macros and with ``--figures`` its figures
  (the reference's figures/make_figures.py).

The reference's ``kernels`` command prints the source its code generator writes for a DSL;
the hip kernels are fixed HIP templates, so there is no such command.

``run`` reports what the device measured: HIP events on the launch stream around the timed
steps, the float64 mass before and after them (a device reduction), and the peak speed of
the final state.  ``--json`` prints that summary as one JSON object, ``--out`` saves the final
``rho`` and ``u`` as npz.
"""

from __future__ import annotations

import argparse
import json
import sys
from typing import Any, Callable

#: run's cases: name -> builder of the Problem from the parsed flags
CASE_BUILDERS: dict[str, Callable[[argparse.Namespace], Any]] = {}


def _case(name: str):
    def register(fn):
        CASE_BUILDERS[name] = fn
        return fn
    return register


@_case("cavity")
def _cavity(a: argparse.Namespace):
    from .problem import Problem

    return Problem.cavity(a.n, re=a.re, u_lid=a.u_lid, collision=a.collision)


@_case("cavity3d")
def _cavity3d(a: argparse.Namespace):
    from .problem import Problem

    return Problem.cavity(a.n, re=a.re, u_lid=a.u_lid, collision=a.collision, dim=3)


@_case("karman")
def _karman(a: argparse.Namespace):
    from .validation.karman import schafer_turek_problem

    return schafer_turek_problem(d=a.d, collision=a.collision)


def run_case(a: argparse.Namespace) -> tuple[dict, Any]:
    """Build the case, make the solver (validation.make_solver: the named entry, or the
    first hip variant that supports the problem), take ``step_multiple`` warm-up steps, then
    time ``steps`` steps (rounded down to the step multiple) with HIP events.  Returns the
    summary and the solver, whose state is then ``steps_total`` steps from the start."""
    import torch

    from .validation import make_solver

    problem = CASE_BUILDERS[a.case](a)
    solver = make_solver(problem, a.backend, a.variant)
    m = solver.step_multiple
    steps = a.steps - a.steps % m
    mass0 = solver.reduce("mass")
    solver.step(m)
    solver.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    if steps:
        solver.step(steps)
    stop.record()
    stop.synchronize()
    seconds = start.elapsed_time(stop) / 1e3
    mass1 = solver.reduce("mass")
    _, u = solver.macroscopic_device()
    speed = float(u.square().sum(dim=0).sqrt().max())
    mlups = problem.n_nodes * steps / seconds / 1e6 if steps and seconds > 0 else 0.0
    summary = {
        "entry": solver.name, "lattice": problem.lattice.name, "size": list(problem.shape),
        "collision": problem.collision, "tau": problem.tau, "storage": solver.storage,
        "steps": steps, "steps_total": steps + m, "seconds": seconds, "MLUPS": mlups,
        "GBs": mlups * solver.bytes_per_node / 1e3,
        "mass": mass1, "mass_change": (mass1 - mass0) / mass0, "max_speed": speed,
    }
    return summary, solver


def _cmd_run(a: argparse.Namespace) -> int:
    summary, solver = run_case(a)
    if a.json:
        print(json.dumps(summary))
    else:
        s = summary
        print(f"{s['entry']} {s['lattice']} {'x'.join(map(str, s['size']))} {s['collision']} "
              f"tau={s['tau']:.4f}: {s['steps']} steps in {s['seconds']:.3f} s on the device, "
              f"{s['MLUPS']:.1f} MLUPS ({s['GBs']:.0f} GB/s by the 2q*itemsize model); "
              f"mass change {s['mass_change']:.2e}, max |u| {s['max_speed']:.4f}")
    if a.out:
        import numpy as np

        rho, u = solver.macroscopic()
        np.savez_compressed(a.out, rho=rho, u=u, steps=summary["steps_total"])
        if not a.json:
            print(f"saved rho, u to {a.out}")
    return 0


def _cmd_backends(_: argparse.Namespace) -> int:
    """One row per registry entry; status() never raises, so neither does this."""
    from . import backends as B

    rows = [("entry", "runs here", "storage", "layout", "B/node D2Q9", "notes")]
    for backend, variant, ok, why in B.status():
        try:
            cls = B.get(backend, variant)
        except Exception as e:  # noqa: BLE001 - a broken entry is reported, not raised
            rows.append((f"{backend}/{variant}", "no", "?", "?", "?", str(e)))
            continue
        storage = getattr(cls, "storage", "?")
        layout = "AA in place" if getattr(cls, "aa", False) else "two buffers"
        item = 2 if storage in ("f16", "bf16") else 4
        rows.append((f"{backend}/{variant}", "yes" if ok else "no", storage, layout,
                     str(2 * 9 * item), cls.description if ok else why))
    width = [max(len(r[k]) for r in rows) for k in range(5)]
    for r in rows:
        print("  ".join(c.ljust(w) for c, w in zip(r[:5], width)) + "  " + r[5])
    return 0


def _cmd_validate(a: argparse.Namespace) -> int:
    from .validation.suite import run_suite

    run_suite(quick=a.quick, cases=a.cases, variant=a.variant, out=a.out,
              vorticity_out=a.vorticity_out)
    return 0


def _cmd_bench(a: argparse.Namespace) -> int:
    from .benchmark import run

    plan = next((name for name in ("full", "dev") if getattr(a, name)), "quick")
    run(plan, out=a.out)
    return 0


def _cmd_report(a: argparse.Namespace) -> int:
    from pathlib import Path

    from . import report

    records = {k: report.load(v) if v else None
               for k, v in (("bench", a.bench), ("validation", a.validation),
                            ("counterparts", a.counterparts))}
    if not any(records.values()):
        raise SystemExit("report: give --bench, --validation and/or --counterparts")
    if a.tex or a.figures:
        from . import paper

        if a.tex:
            src = Path(a.bench).name if a.bench else "none"
            Path(a.tex).write_text(paper.results_tex(records["bench"], records["validation"],
                                                     records["counterparts"], source=src))
            print(f"wrote {a.tex}")
        if a.figures:
            vort = None
            if a.vorticity:
                import numpy as np

                with np.load(a.vorticity) as z:
                    vort = {"vorticity": z["vorticity"], "solid": z["solid"]}
            for p in paper.figures(a.figures, records["bench"], records["validation"],
                                   records["counterparts"], vort):
                print(f"wrote {p}")
        if not (a.out or a.readme):
            return 0
    if a.readme:
        path = Path(a.readme)
        blocks = report.blocks_for(records["bench"], records["validation"])
        path.write_text(report.fill_blocks(path.read_text(), blocks))
        print(f"updated {path}")
    text = report.render(records["bench"], records["validation"])
    if a.out:
        Path(a.out).write_text(text)
        print(f"wrote {a.out}")
    elif not a.readme:
        print(text)
    return 0


def _add_run(sp: argparse.ArgumentParser) -> None:
    sp.add_argument("case", choices=sorted(CASE_BUILDERS))
    sp.add_argument("--n", type=int, default=256, help="cavity: nodes per side")
    sp.add_argument("--d", type=int, default=20, help="karman: nodes per cylinder diameter")
    sp.add_argument("--re", type=float, default=100.0)
    sp.add_argument("--u-lid", type=float, default=0.1)
    sp.add_argument("--collision", default="bgk", choices=["bgk", "trt", "mrt"])
    sp.add_argument("--backend", default=None, help="registry backend (hip)")
    sp.add_argument("--variant", default=None, help="e.g. fused, fused-aa, fused-fp16")
    sp.add_argument("--steps", type=int, default=10000)
    sp.add_argument("--json", action="store_true", help="print the summary as JSON")
    sp.add_argument("--out", default=None, help="save the final rho and u (npz)")


def _add_validate(sp: argparse.ArgumentParser) -> None:
    sp.add_argument("--quick", action="store_true", help="small grids")
    sp.add_argument("--cases", nargs="*", default=None,
                    help="subset of: ghia shear poiseuille karman tgv")
    sp.add_argument("--variant", default=None, help="run every case on this hip variant")
    sp.add_argument("--out", default=None, help="write the record (JSON)")
    sp.add_argument("--vorticity-out", default=None, help="write the Karman vorticity (npz)")


def _add_bench(sp: argparse.ArgumentParser) -> None:
    g = sp.add_mutually_exclusive_group()
    for plan in ("quick", "dev", "full"):
        g.add_argument(f"--{plan}", action="store_true")
    sp.add_argument("--out", default=None)


def _add_report(sp: argparse.ArgumentParser) -> None:
    sp.add_argument("--bench", default=None, help="bench --out record or bench.py output")
    sp.add_argument("--validation", default=None, help="validate --out record")
    sp.add_argument("--out", default=None, help="write the report here (default: stdout)")
    sp.add_argument("--readme", default=None,
                    help="refresh the <!-- pylbm_amd:NAME --> blocks of this Markdown file")
    sp.add_argument("--counterparts", default=None, help="head-to-head record (JSON)")
    sp.add_argument("--tex", default=None, help="write the paper's results.tex macros here")
    sp.add_argument("--figures", default=None,
                    help="draw the paper's figures (pdf and png) into this folder")
    sp.add_argument("--vorticity", default=None,
                    help="validate --vorticity-out file for the cylinder figure")


#: command -> (help, argument builder, handler)
COMMANDS = {
    "backends": ("list the registry entries and whether they run here", None, _cmd_backends),
    "run": ("run one case on the device and report its throughput", _add_run, _cmd_run),
    "validate": ("run the validation suite", _add_validate, _cmd_validate),
    "bench": ("benchmark every hip variant (cavity, Re = 1000)", _add_bench, _cmd_bench),
    "report": ("Markdown tables, paper macros and figures from bench / validate records",
               _add_report, _cmd_report),
}


def parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(prog="pylbm_amd",
                                 description="fused lattice-Boltzmann step on gfx950 (hip)")
    sub = ap.add_subparsers(dest="cmd", required=True)
    for name, (help_text, build, handler) in COMMANDS.items():
        sp = sub.add_parser(name, help=help_text)
        if build:
            build(sp)
        sp.set_defaults(handler=handler)
    return ap


def main(argv: list[str] | None = None) -> int:
    a = parser().parse_args(argv)
    return a.handler(a)


if __name__ == "__main__":  # pragma: no cover
    sys.exit(main(sys.argv[1:]))
