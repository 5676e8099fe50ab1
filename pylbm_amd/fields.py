"""Seeded synthetic fields for the differentiable runs (numpy, float64, array layout)."""

from __future__ import annotations

import numpy as np


def smooth_random_field(shape: tuple[int, ...], modes: int, amplitude: float,
                        seed: int) -> np.ndarray:
    """(d, *array_shape) field: a sum of ``modes`` random low Fourier modes per component.

    ``shape`` is the problem shape (nx, ny[, nz]); wavenumbers are integers in 1..3 per axis
    so the field is periodic; the result is scaled to max |component| = amplitude.
    """
    rng = np.random.default_rng(seed)
    d = len(shape)
    grids = np.meshgrid(*[2 * np.pi * np.arange(n) / n for n in reversed(shape)],
                        indexing="ij")
    coord = [grids[d - 1 - a] for a in range(d)]  # problem axis a
    out = np.zeros((d,) + tuple(reversed(shape)))
    for c in range(d):
        for _ in range(modes):
            k = rng.integers(1, 4, size=d)
            ph = rng.uniform(0, 2 * np.pi)
            out[c] += rng.standard_normal() * np.sin(sum(k[a] * coord[a] for a in range(d)) + ph)
    return amplitude * out / np.abs(out).max()


def divergence_free_field(n: int, modes: int, amplitude: float, seed: int) -> np.ndarray:
    """(3, n, n, n) periodic divergence-free velocity: the curl of a random vector
    potential built from ``modes`` low Fourier modes per component (spectral derivative),
    scaled to max |u| = amplitude."""
    rng = np.random.default_rng(seed)
    k = 2 * np.pi / n
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    X = (x, y, z)
    A = np.zeros((3, n, n, n))
    dA = np.zeros((3, 3, n, n, n))  # dA[c, a] = d A_c / d x_a
    for c in range(3):
        for _ in range(modes):
            kv = rng.integers(-2, 3, size=3)
            if not kv.any():
                kv[0] = 1
            amp, ph = rng.standard_normal(), rng.uniform(0, 2 * np.pi)
            arg = k * (kv[0] * X[0] + kv[1] * X[1] + kv[2] * X[2]) + ph
            A[c] += amp * np.sin(arg)
            for a in range(3):
                dA[c, a] += amp * k * kv[a] * np.cos(arg)
    u = np.stack([dA[2, 1] - dA[1, 2], dA[0, 2] - dA[2, 0], dA[1, 0] - dA[0, 1]])
    return amplitude * u / np.abs(u).max()
