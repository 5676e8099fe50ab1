"""Write the golden vectors that pin the float64 oracle to the reference's own NumPy solver.

Run in a container that has the reference checkout (read-only) at /root/reference:

    PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden.py

The reference is imported from /root/reference/src only by this script; the outputs are
data (inputs and the reference's outputs) saved as compressed float64 ``.npz`` files next
to this script.  What is generated:

* ``equivalence.npz``: the seven box cases of tests/test_equivalence.py:9-20 advanced 100
  steps with the reference NumpySolver (pylbm/backends/numpy_backend.py:12-36).
* ``cavity128.npz``: the 128x128 BGK cavity, tau = 0.884 (Re = 100, U = 0.1), after 1000
  steps; stored as every 4th node plus the full lid row, the full left column and global
  sums (the full state is 1.2 MB).
* ``units.npz``: single ``stream`` and ``collide`` calls (pylbm/core.py:88-127,203-247) on
  seeded random populations for tiny grids: walls, lid corners, mixed periodicity, uniform
  force, BGK/TRT/MRT, and MRT with every non-conserved rate overridden (the reference does
  this by monkeypatching ``mrt_rate_tuple``, tests/test_lattice.py:45-46).
* ``geometry.npz`` (``--only geometry`` writes just this file): general geometry
  (pylbm/core.py:56-148).  Single ``stream`` calls on seeded random populations for tiny
  grids with random solid nodes, a velocity inlet, the pressure outlet and a lid; the
  Schaefer-Turek channel of tests/test_equivalence.py:81-92 (d = 6, 60 steps, TRT as in
  that test, BGK and MRT besides) with its inputs; 2D and 3D obstacle trajectories with
  force and lid.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

REF_SRC = Path("/root/reference/src")
OUT = Path(__file__).resolve().parent


def geometry() -> None:
    from pylbm.backends.numpy_backend import NumpySolver
    from pylbm.core import Ops, make_geometry, stream
    from pylbm.problem import Problem
    from pylbm.validation.karman import schafer_turek_problem

    rng = np.random.default_rng(20261204)
    out = {}

    def rand_solid(shape_arr, frac):
        return rng.random(shape_arr) < frac

    stream_cases = {
        "gs2_all": dict(shape=(7, 6), u_lid=0.07, inlet=True, outlet=True),
        "gs2_perx": dict(shape=(6, 5), periodic=(True, False), inlet=True, outlet=True),
        "gs2_pery": dict(shape=(7, 5), periodic=(False, True), inlet=True, outlet=True),
        "gs2_inlet_lid": dict(shape=(6, 4), u_lid=0.05, inlet=True),
        "gs3_solid_lid": dict(shape=(5, 4, 6), periodic=(True, False, False), u_lid=0.05),
        "gs3_solid_per": dict(shape=(4, 5, 3), periodic=(True, True, True)),
    }
    for name, kw in stream_cases.items():
        kw = dict(kw)
        shape = kw.pop("shape")
        ashape = tuple(reversed(shape))
        solid = rand_solid(ashape, 0.25)
        inlet = 0.05 * rng.random(ashape[-2]) if kw.pop("inlet", False) else None
        pr = Problem(shape, solid=solid, inlet=inlet, **kw)
        q = pr.lattice.q
        f0 = pr.lattice.w.reshape((-1,) + (1,) * pr.dim) * (
            1.0 + 0.1 * rng.standard_normal((q,) + pr.array_shape))
        res = np.stack(stream(f0, make_geometry(pr, np.asarray), Ops(np)))
        out[f"{name}/in"] = f0
        out[f"{name}/out"] = res
        out[f"{name}/solid"] = solid
        if inlet is not None:
            out[f"{name}/inlet"] = inlet
    for coll in ("trt", "bgk", "mrt"):
        pr = schafer_turek_problem(d=6, u_mean=0.05, collision=coll)
        s = NumpySolver(pr)
        s.step(60)
        out[f"karman6_{coll}/f"] = s.populations()
        if coll == "trt":
            out["karman6/solid"] = pr.solid
            out["karman6/inlet"] = pr.inlet
            out["karman6/init_u"] = pr.init_u
            out["karman6/tau"] = np.float64(pr.tau)
            out["karman6/shape"] = np.array(pr.shape)
        print(f"karman6 {coll}: {out[f'karman6_{coll}/f'].shape}")
    # 2D obstacle channel, MRT + force; 3D obstacle box, TRT + force + lid
    solid2 = np.zeros((16, 24), dtype=bool)
    solid2[5:10, 6:9] = True
    solid2[2, 15] = True
    pr = Problem((24, 16), tau=0.6, collision="mrt", periodic=(True, False),
                 force=(1e-5, -2e-6), solid=solid2)
    s = NumpySolver(pr)
    s.step(60)
    out["geo2d_mrt_force/f"] = s.populations()
    out["geo2d_mrt_force/solid"] = solid2
    solid3 = np.zeros((8, 10, 12), dtype=bool)
    solid3[2:5, 3:6, 4:8] = True
    solid3[6, 8, 1] = True
    pr = Problem((12, 10, 8), tau=0.7, collision="trt", periodic=(True, False, False),
                 force=(1e-5, 0.0, 2e-6), u_lid=0.05, solid=solid3)
    s = NumpySolver(pr)
    s.step(60)
    out["geo3d_trt_lid/f"] = s.populations()
    out["geo3d_trt_lid/solid"] = solid3
    np.savez_compressed(OUT / "geometry.npz", **out)
    print(f"geometry: {len(out)} arrays")


def main() -> None:
    sys.dont_write_bytecode = True
    sys.path.insert(0, str(REF_SRC))
    if "--only" in sys.argv and sys.argv[sys.argv.index("--only") + 1] == "geometry":
        geometry()
        return
    geometry()
    from pylbm.backends.numpy_backend import NumpySolver
    from pylbm.core import Ops, collide, make_geometry, stream
    from pylbm.problem import Problem

    # ------------------------------------------------------------------ equivalence cases
    cases = {
        "cavity2d-bgk": dict(shape=(40, 28), tau=0.7, u_lid=0.1),
        "cavity2d-trt": dict(shape=(40, 28), tau=0.7, u_lid=0.1, collision="trt"),
        "cavity2d-mrt": dict(shape=(40, 28), tau=0.7, u_lid=0.1, collision="mrt"),
        "channel2d-force": dict(shape=(40, 28), tau=0.7, periodic=(True, False),
                                force=(1e-5, 2e-6)),
        "periodic2d-mrt-force": dict(shape=(32, 24), tau=0.6, periodic=(True, True),
                                     force=(1e-5, -3e-6), collision="mrt"),
        "cavity3d-bgk": dict(shape=(20, 14, 12), tau=0.7, u_lid=0.1),
        "periodic3d-trt": dict(shape=(16, 12, 10), tau=0.7, periodic=(True, True, False),
                               force=(1e-5, 1e-6, 0.0), collision="trt"),
    }
    eq = {}
    for name, kw in cases.items():
        s = NumpySolver(Problem(**kw))
        s.step(100)
        eq[name] = s.populations()
        print(f"equivalence {name}: {eq[name].shape}")
    np.savez_compressed(OUT / "equivalence.npz", **eq)

    # ------------------------------------------------------------------ config-1 cavity
    p = Problem.cavity(128, re=100.0, u_lid=0.1)
    assert abs(p.tau - 0.884) < 1e-12
    s = NumpySolver(p)
    s.step(1000)
    f = s.populations()
    np.savez_compressed(
        OUT / "cavity128.npz",
        sub=f[:, ::4, ::4], lid_row=f[:, -1, :], left_col=f[:, :, 0],
        sum=f.sum(axis=(1, 2)), sumsq=(f ** 2).sum(axis=(1, 2)), tau=p.tau, steps=1000)
    print("cavity128 done")

    # ------------------------------------------------------------------ unit vectors
    rng = np.random.default_rng(20261203)
    units = {}
    stream_cases = {
        "s2_walls_lid": dict(shape=(5, 4), u_lid=0.07),
        "s2_perx_lid": dict(shape=(5, 4), periodic=(True, False), u_lid=0.05),
        "s2_pery": dict(shape=(6, 5), periodic=(False, True)),
        "s2_periodic": dict(shape=(5, 3), periodic=(True, True)),
        "s3_walls_lid": dict(shape=(4, 3, 5), u_lid=0.06),
        "s3_mixed_lid": dict(shape=(4, 5, 3), periodic=(True, False, False), u_lid=0.04),
        "s3_perz": dict(shape=(3, 4, 5), periodic=(False, False, True)),
    }
    for name, kw in stream_cases.items():
        pr = Problem(**kw)
        q = pr.lattice.q
        f0 = pr.lattice.w.reshape((-1,) + (1,) * pr.dim) * (
            1.0 + 0.1 * rng.standard_normal((q,) + pr.array_shape))
        out = np.stack(stream(f0, make_geometry(pr, np.asarray), Ops(np)))
        units[f"{name}/in"] = f0
        units[f"{name}/out"] = out
    collide_cases = {
        "c2_bgk": dict(shape=(7, 6), tau=0.8),
        "c2_bgk_force": dict(shape=(7, 6), tau=0.8, force=(1e-4, -2e-4)),
        "c2_trt_force": dict(shape=(7, 6), tau=0.6, collision="trt", force=(3e-4, 1e-4)),
        "c2_mrt": dict(shape=(7, 6), tau=0.7, collision="mrt"),
        "c2_mrt_force": dict(shape=(7, 6), tau=0.6, collision="mrt", force=(-1e-4, 2e-4),
                             s_e=1.2, s_eps=1.2),
        "c3_bgk_force": dict(shape=(4, 5, 3), tau=0.7, force=(1e-4, 2e-5, -5e-5)),
        "c3_trt": dict(shape=(4, 5, 3), tau=0.55, collision="trt"),
        "c3_trt_force": dict(shape=(4, 5, 3), tau=0.9, collision="trt",
                             force=(-1e-4, 3e-5, 2e-4)),
    }
    for name, kw in collide_cases.items():
        pr = Problem(**kw)
        q = pr.lattice.q
        f0 = pr.lattice.w.reshape((-1,) + (1,) * pr.dim) * (
            1.0 + 0.05 * rng.standard_normal((q,) + pr.array_shape))
        units[f"{name}/in"] = f0
        units[f"{name}/out"] = collide(list(f0), pr, Ops(np))
    # MRT with s_q (and every other non-conserved rate) set to 1.2 except shear
    pr = Problem((7, 6), tau=0.6, collision="mrt", force=(2e-4, -1e-4), s_e=1.2, s_eps=1.2)
    pr.mrt_rate_tuple = lambda: (0.0, 1.2, 1.2, 0.0, 1.2, 0.0, 1.2, 1 / 0.6, 1 / 0.6)
    f0 = pr.lattice.w.reshape(-1, 1, 1) * (1.0 + 0.05 * rng.standard_normal((9, 6, 7)))
    units["c2_mrt_sq12/in"] = f0
    units["c2_mrt_sq12/out"] = collide(list(f0), pr, Ops(np))
    np.savez_compressed(OUT / "units.npz", **units)
    print(f"units: {len(units) // 2} cases")


if __name__ == "__main__":
    main()
