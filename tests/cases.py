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
TOL = {"f32": 3e-6, "f16": 1e-4, "bf16": 6e-4}


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
