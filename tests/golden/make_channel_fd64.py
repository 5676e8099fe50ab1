"""Write float64 finite-difference gradients of the differentiable channel at its full size.

    python tests/golden/make_channel_fd64.py        # about 3 minutes on 8 cores

The channel: D2Q9 BGK 1024 x 256, periodic in x with walls in y, tau = 0.8, per-cell Guo
force 1e-6 + 1e-7 N(0,1) (seed 20261203, rounded to fp32 as the kernels read it), loss
0.5 sum |u(T) - u_target|^2 after T = 2000 steps with the smooth 4-mode target of seed
20261204.  The float64 C oracle (oracle/lbm_ref_c.c, itself pinned to the reference's
NumPy solver) runs the forward; the script writes the loss, dL/dtau by central differences
at two steps (1e-4 and 1e-3 of tau) and <dL/dF, V> for a seeded normal direction V
(numpy seed 11) at two steps (1e-2 and 1e-3 of max |F|).  The two steps of each derivative
agree to a few parts in 1e-6 or better, so the values stand for the float64 model's
derivatives.  Output: ``channel_fd64.json`` next to this script (data only).
"""

from __future__ import annotations

import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import lbm_ref as R  # noqa: E402
from oracle import lbm_ref_c as RC  # noqa: E402
from tests.test_adjoint import _channel_setup  # noqa: E402

NX, NY, T, TAU, SEED, V_SEED = 1024, 256, 2000, 0.8, 20261203, 11


def main() -> None:
    RC.build()
    ff, ut = _channel_setup(NX, NY, seed=SEED)
    kw = dict(shape=(NX, NY), periodic=(True, False), collision="bgk")

    def loss(tau: float, field: np.ndarray) -> float:
        c = R.Case(**kw, tau=tau, force_field=field)
        t0 = time.time()
        f = RC.run(c, T)
        _, u = R.macroscopic(f, c)
        val = 0.5 * float(((u - ut) ** 2).sum())
        print(f"  L(tau={tau:.6f}) = {val!r} ({time.time() - t0:.1f} s)", flush=True)
        return val

    out = {"shape": [NX, NY], "steps": T, "tau": TAU, "seed": SEED, "v_seed": V_SEED,
           "loss": loss(TAU, ff), "dL_dtau": {}, "dL_dF_dot_V": {}}
    for rel in (1e-4, 1e-3):
        h = rel * TAU
        out["dL_dtau"][f"{rel:g}"] = (loss(TAU + h, ff) - loss(TAU - h, ff)) / (2 * h)
    V = np.random.default_rng(V_SEED).standard_normal(ff.shape)
    for rel in (1e-2, 1e-3):
        h = rel * float(np.abs(ff).max())
        out["dL_dF_dot_V"][f"{rel:g}"] = (loss(TAU, ff + h * V) - loss(TAU, ff - h * V)) / (2 * h)
    path = Path(__file__).resolve().parent / "channel_fd64.json"
    path.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {path}: {json.dumps(out)}")


if __name__ == "__main__":
    main()
