"""Test cases shared by the oracle and HIP parity tests.

EQUIVALENCE are the seven box problems of the reference's backend-equivalence test
(tests/test_equivalence.py:9-20); their float64 states after 100 steps are stored in
tests/golden/equivalence.npz, written from the reference NumPy solver.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"

EQUIVALENCE = {
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

#: reference tolerances (tests/test_equivalence.py:21)
TOL = {"f32": 3e-6, "f32s": 3e-6, "f16": 1e-4, "bf16": 6e-4}


def golden(name: str) -> dict:
    with np.load(GOLDEN / name, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def shear_wave_fields(n: int, seed: int = 0, amp: float = 0.01, noise: float = 1e-3):
    """Periodic shear wave u_x = amp sin(2 pi y / n), rho = 1 + noise U(-1, 1) per cell."""
    rng = np.random.default_rng(seed)
    y = np.arange(n)
    u = np.zeros((2, n, n))
    u[0] = amp * np.sin(2 * np.pi * y / n)[:, None]
    rho = 1.0 + noise * rng.uniform(-1.0, 1.0, size=(n, n))
    return rho, u


def tgv3d_fields(n: int, amp: float = 0.02):
    """u = amp (sin x cos y cos z, -cos x sin y cos z, 0) on x_k = 2 pi k / n, rho = 1."""
    k = 2 * np.pi * np.arange(n) / n
    z, y, x = np.meshgrid(k, k, k, indexing="ij")
    u = np.zeros((3, n, n, n))
    u[0] = amp * np.sin(x) * np.cos(y) * np.cos(z)
    u[1] = -amp * np.cos(x) * np.sin(y) * np.cos(z)
    return np.ones((n, n, n)), u


#: general-geometry stream cases of tests/golden/geometry.npz (solid / inlet stored there)
GEO_STREAM = {
    "gs2_all": dict(shape=(7, 6), u_lid=0.07, outlet=True),
    "gs2_perx": dict(shape=(6, 5), periodic=(True, False), outlet=True),
    "gs2_pery": dict(shape=(7, 5), periodic=(False, True), outlet=True),
    "gs2_inlet_lid": dict(shape=(6, 4), u_lid=0.05),
    "gs3_solid_lid": dict(shape=(5, 4, 6), periodic=(True, False, False), u_lid=0.05),
    "gs3_solid_per": dict(shape=(4, 5, 3), periodic=(True, True, True)),
}

#: general-geometry trajectories of tests/golden/geometry.npz (60 steps each)
GEO_TRAJ = {
    "geo2d_mrt_force": dict(shape=(24, 16), tau=0.6, collision="mrt", periodic=(True, False),
                            force=(1e-5, -2e-6)),
    "geo3d_trt_lid": dict(shape=(12, 10, 8), tau=0.7, collision="trt",
                          periodic=(True, False, False), force=(1e-5, 0.0, 2e-6), u_lid=0.05),
}


def geo_stream_kwargs(g: dict, name: str) -> dict:
    kw = dict(GEO_STREAM[name], solid=g[f"{name}/solid"])
    if f"{name}/inlet" in g:
        kw["inlet"] = g[f"{name}/inlet"]
    return kw


def karman6_kwargs(g: dict, collision: str) -> dict:
    """The Schaefer-Turek d = 6 channel of geometry.npz as problem keyword arguments."""
    return dict(shape=tuple(int(v) for v in g["karman6/shape"]), tau=float(g["karman6/tau"]),
                collision=collision, periodic=(False, False), solid=g["karman6/solid"],
                inlet=g["karman6/inlet"], outlet=True, init_u=g["karman6/init_u"])
