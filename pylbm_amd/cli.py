"""Command line interface: ``python -m pylbm_amd run | validate | bench | report | backends``.

Mirrors the reference's ``pylbm`` console script (pylbm/cli.py:98-144) for the hip backend:
``run cavity --backend hip --variant fused|fused-aa|fused-fp16|fused-bf16`` runs the box
kernels, ``run karman`` the obstacle / inlet / outlet kernel (variant ``gather``).
``validate`` runs the validation suite and ``report`` turns bench / validate records into
Markdown tables.  (The reference's ``kernels`` subcommand prints DSL source; there is none.)
"""

from __future__ import annotations

import argparse
import sys
import time


def _cmd_backends(_: argparse.Namespace) -> None:
    from . import backends as B

    for b, v, ok, why in B.status():
        desc = B.get(b, v).description if ok else why
        print(f"{'ok ' if ok else '-- '} {b + '/' + v:28s} {desc}")


def _cmd_run(a: argparse.Namespace) -> None:
    from .problem import Problem
    from .validation import make_solver

    if a.case == "cavity":
        p = Problem.cavity(a.n, re=a.re, u_lid=a.u_lid, collision=a.collision)
    elif a.case == "cavity3d":
        p = Problem.cavity(a.n, re=a.re, u_lid=a.u_lid, collision=a.collision, dim=3)
    elif a.case == "karman":
        from .validation.karman import schafer_turek_problem

        p = schafer_turek_problem(d=a.d, collision=a.collision)
        if a.variant == "fused" and a.backend == "hip":
            a.variant = "gather"
    else:
        raise SystemExit(f"unknown case {a.case}")
    s = make_solver(p, a.backend, a.variant)
    steps = a.steps - a.steps % s.step_multiple
    print(f"{s.name}: {p.lattice.name} {'x'.join(map(str, p.shape))}, tau={p.tau:.4f}, "
          f"{steps} steps")
    s.step(s.step_multiple)
    s.synchronize()
    t0 = time.perf_counter()
    s.step(steps)
    s.synchronize()
    dt = time.perf_counter() - t0
    print(f"{p.n_nodes * steps / dt / 1e6:.1f} MLUPS ({dt:.2f} s, not a benchmark: use bench)")
    if a.out:
        import numpy as np

        rho, u = s.macroscopic()
        np.savez_compressed(a.out, rho=rho, u=u)
        print(f"wrote {a.out}")


def _cmd_validate(a: argparse.Namespace) -> None:
    from .validation.suite import run_suite

    run_suite(quick=a.quick, cases=a.cases, variant=a.variant, out=a.out,
              vorticity_out=a.vorticity_out)


def _cmd_bench(a: argparse.Namespace) -> None:
    from .benchmark import run

    run("full" if a.full else "dev" if a.dev else "quick", out=a.out)


def _cmd_report(a: argparse.Namespace) -> None:
    from pathlib import Path

    from . import report

    bench = report.load(a.bench) if a.bench else None
    val = report.load(a.validation) if a.validation else None
    if bench is None and val is None:
        raise SystemExit("report: give --bench and/or --validation")
    if a.readme:
        path = Path(a.readme)
        path.write_text(report.fill_blocks(path.read_text(), report.blocks_for(bench, val)))
        print(f"updated {path}")
    text = report.render(bench, val)
    if a.out:
        Path(a.out).write_text(text)
        print(f"wrote {a.out}")
    elif not a.readme:
        print(text)


def main(argv: list[str] | None = None) -> None:
    ap = argparse.ArgumentParser(prog="pylbm_amd", description="gfx950 fused LBM (hip backend)")
    sub = ap.add_subparsers(dest="cmd", required=True)
    sp = sub.add_parser("backends", help="list backends and whether they run here")
    sp.set_defaults(func=_cmd_backends)
    sp = sub.add_parser("run", help="run a case and report throughput")
    sp.add_argument("case", choices=["cavity", "cavity3d", "karman"])
    sp.add_argument("--d", type=int, default=20, help="karman: nodes per cylinder diameter")
    sp.add_argument("--n", type=int, default=256)
    sp.add_argument("--re", type=float, default=100.0)
    sp.add_argument("--u-lid", type=float, default=0.1)
    sp.add_argument("--collision", default="bgk", choices=["bgk", "trt", "mrt"])
    sp.add_argument("--backend", default="hip")
    sp.add_argument("--variant", default="fused")
    sp.add_argument("--steps", type=int, default=10000)
    sp.add_argument("--out", default=None, help="write rho, u (npz)")
    sp.set_defaults(func=_cmd_run)
    sp = sub.add_parser("validate", help="run the validation suite on the hip backend")
    sp.add_argument("--quick", action="store_true", help="small grids")
    sp.add_argument("--cases", nargs="*", default=None,
                    help="subset of: ghia shear poiseuille karman tgv")
    sp.add_argument("--variant", default=None, help="run every case on this hip variant")
    sp.add_argument("--out", default=None, help="write the record (JSON)")
    sp.add_argument("--vorticity-out", default=None, help="write the Karman vorticity (npz)")
    sp.set_defaults(func=_cmd_validate)
    sp = sub.add_parser("bench", help="benchmark every hip variant (cavity, Re=1000)")
    g = sp.add_mutually_exclusive_group()
    g.add_argument("--quick", action="store_true")
    g.add_argument("--dev", action="store_true")
    g.add_argument("--full", action="store_true")
    sp.add_argument("--out", default=None)
    sp.set_defaults(func=_cmd_bench)
    sp = sub.add_parser("report", help="Markdown tables from bench / validate JSON")
    sp.add_argument("--bench", default=None, help="bench --out record or bench.py output")
    sp.add_argument("--validation", default=None, help="validate --out record")
    sp.add_argument("--out", default=None, help="write the report here (default: stdout)")
    sp.add_argument("--readme", default=None,
                    help="refresh the <!-- pylbm_amd:NAME --> blocks of this Markdown file")
    sp.set_defaults(func=_cmd_report)
    a = ap.parse_args(argv)
    a.func(a)


if __name__ == "__main__":  # pragma: no cover
    main(sys.argv[1:])
