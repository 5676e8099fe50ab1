"""Karman vortex street: Schaefer & Turek (1996) benchmark 2D-2 on the hip backend.

Mirrors pylbm/validation/karman.py:1-139: channel 2.2 x 0.41, cylinder of diameter
D = 0.1 centred at (0.2, 0.2) as a staircase of solid nodes, parabolic inflow with mean
velocity U (Ladd inlet), rho = 1 pressure outlet, Re = U D / nu = 100.  Drag and lift come
from the momentum exchange over the fluid-solid links, evaluated on the device from the
decoded state; the Strouhal number from the dominant lift frequency.  The benchmark's
reference intervals (Table 4 of the paper) are kept in REFERENCE.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from ..problem import Problem
from . import make_solver

REFERENCE = {"St": (0.2950, 0.3050), "cd_max": (3.2200, 3.2400), "cl_max": (0.9900, 1.0100)}


def schafer_turek_problem(d: int = 20, u_mean: float = 0.05, collision: str = "trt") -> Problem:
    """Benchmark 2D-2 with ``d`` nodes per cylinder diameter (pylbm/validation/karman.py:21-43).

    The geometry, inflow and initial field are the benchmark's definition, so they are the
    same numbers as the reference's: nodes at cell centres, walls half-way at y = 0 and
    y = ny, parabolic inflow 4 U_max y (h - y) / h^2 with U_max = 1.5 U, an initial velocity
    equal to the inflow profile (zero inside the cylinder) plus a small transverse Gaussian
    kick that starts the shedding.
    """
    dx = 0.1 / d
    nx, ny = round(2.2 / dx), round(0.41 / dx)
    xc = (np.arange(nx) + 0.5) * dx
    yc = (np.arange(ny) + 0.5) * dx
    X, Y = np.meshgrid(xc, yc)
    solid = (X - 0.2) ** 2 + (Y - 0.2) ** 2 <= 0.05 ** 2
    u_max = 1.5 * u_mean
    yy = np.arange(ny) + 0.5
    inlet = 4.0 * u_max * yy * (ny - yy) / ny ** 2
    nu = u_mean * d / 100.0
    init_u = np.zeros((2, ny, nx))
    init_u[0] = inlet[:, None]
    init_u[0][solid] = 0.0
    init_u[1] = 0.05 * u_mean * np.exp(-((X - 0.3) ** 2 + (Y - 0.2) ** 2) / 0.05 ** 2)
    return Problem((nx, ny), tau=3.0 * nu + 0.5, collision=collision, periodic=(False, False),
                   solid=solid, inlet=inlet, outlet=True, init_u=init_u,
                   meta={"d": d, "u_mean": u_mean, "dx": dx})


def force_masks(p: Problem) -> np.ndarray:
    """(q, ny, nx) bool: fluid nodes x whose neighbour x + c_i is solid
    (pylbm/validation/karman.py:46-55)."""
    lat = p.lattice
    fluid = ~p.solid
    out = np.zeros((lat.q,) + p.array_shape, dtype=bool)
    for i in range(lat.q):
        nb = np.roll(p.solid, (-int(lat.c[i, 1]), -int(lat.c[i, 0])), axis=(0, 1))
        out[i] = fluid & nb
    return out


class MomentumExchange:
    """Force on the solid from the stored post-collision state: F = 2 sum_i c_i sum_x f*_i(x)
    over the fluid nodes x with a solid neighbour at x + c_i (the bounce-back of f*_i
    returns -f*_i, pylbm/validation/karman.py:86-97).  Evaluated on the device."""

    def __init__(self, solver: Any) -> None:
        import torch

        p = solver.problem
        self.solver = solver
        dev = solver.device
        m = force_masks(p)
        self.idx = [torch.as_tensor(np.flatnonzero(m[i]), device=dev) for i in range(p.lattice.q)]
        self.c = p.lattice.c.astype(np.float64)

    def __call__(self) -> np.ndarray:
        import torch

        f = self.solver.populations_device().reshape(len(self.idx), -1)
        per_dir = torch.stack([f[i].index_select(0, ix).sum() if ix.numel() else
                               f.new_zeros(()) for i, ix in enumerate(self.idx)])
        return 2.0 * (per_dir.cpu().numpy()[:, None] * self.c).sum(axis=0)


def strouhal(cl: np.ndarray, sample_every: int, d: int, u_mean: float) -> float:
    """Dominant lift frequency from a Hann-windowed, 16x zero-padded FFT with parabolic
    interpolation of the peak, in units of U / D."""
    sig = cl - cl.mean()
    nfft = 16 * len(sig)
    spec = np.abs(np.fft.rfft(sig * np.hanning(len(sig)), nfft))
    k = float(np.argmax(spec[1:]) + 1)
    ki = int(k)
    if 1 <= ki < len(spec) - 1:
        a, b, g = spec[ki - 1], spec[ki], spec[ki + 1]
        k = ki + 0.5 * (a - g) / (a - 2 * b + g)
    return k / (nfft * sample_every) * d / u_mean


def run_karman(d: int = 20, u_mean: float = 0.05, periods: float = 20.0,
               t_transient: float = 60.0, collision: str = "trt", backend: str | None = "hip",
               variant: str | None = "gather", sample_every: int = 10,
               factory: Any = None) -> dict:
    """Run benchmark 2D-2 and measure St, max c_D, max c_L (pylbm/validation/karman.py:58-139).

    ``t_transient`` and ``periods`` are in convective units D / U; the statistics use the
    ``periods`` units after the transient.
    """
    p = schafer_turek_problem(d, u_mean, collision)
    s = make_solver(p, backend, variant, factory)
    force = MomentumExchange(s)
    t_conv = d / u_mean
    n_trans = int(t_transient * t_conv)
    n_meas = int(periods * t_conv)
    m = s.step_multiple
    s.step(n_trans - n_trans % m)
    ts, fs = [], []
    done = 0
    while done < n_meas:
        s.step(sample_every)
        done += sample_every
        ts.append(s.steps_done)
        fs.append(force())
    fs_a = np.array(fs)
    scale = 2.0 / (u_mean ** 2 * d)
    cd, cl = fs_a[:, 0] * scale, fs_a[:, 1] * scale
    st = strouhal(cl, sample_every, d, u_mean)
    rho, u = s.macroscopic()
    vort = np.gradient(u[1], axis=1) - np.gradient(u[0], axis=0)
    vort[p.solid] = np.nan
    return {"d": d, "u_mean": u_mean, "tau": p.tau, "collision": collision,
            "backend": s.name, "shape": list(p.shape), "steps": s.steps_done,
            "St": float(st), "cd_max": float(cd.max()), "cl_max": float(cl.max()),
            "cd_mean": float(cd.mean()), "reference": REFERENCE,
            "t_conv": (np.array(ts) / t_conv).tolist()[::5], "cl": cl.tolist()[::5],
            "cd": cd.tolist()[::5], "_vorticity": vort * d / u_mean, "_solid": p.solid}
