"""pylbm_hip_reduce_noalloc (GPU): the allocation-free reduction equals pylbm_hip_reduce.

The entry keeps its per-workgroup partials in the library's static device array instead of
a stream-ordered allocation.  Through the C ABI, for every reduction kind, storage and
lattice: two calls in a row, the second on a smaller problem (so the second call meets the
first call's partials still in the array), each equal to pylbm_hip_reduce bit for bit.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from pylbm_amd import _lib
from pylbm_amd import backends as B
from pylbm_amd.problem import Problem

pytestmark = pytest.mark.gpu

KINDS = {"mass": _lib.RED_MASS, "kinetic": _lib.RED_KINETIC,
         "maxabs_diff": _lib.RED_MAXABS_DIFF, "loss": _lib.RED_LOSS, "dot": _lib.RED_DOT}


def _states(variant: str, shape: tuple, seed: int):
    dim = len(shape)
    p = Problem(shape, tau=0.7, periodic=(True,) * dim, force=(1e-5,) + (0.0,) * (dim - 1))
    rng = np.random.default_rng(seed)
    a, b = (B.make("hip", variant, p, init="none") for _ in range(2))
    rho = 1 + 0.01 * rng.standard_normal(p.array_shape)
    u = 0.02 * rng.standard_normal((dim,) + p.array_shape)
    for s in (a, b):
        s.init_equilibrium(rho, u)
    b.step(b.step_multiple)
    target = torch.as_tensor(0.01 * rng.standard_normal((dim,) + p.array_shape),
                             device="cuda")
    return a, b, target


def _call(name: str, s, kind: str, other, out) -> None:
    k = KINDS[kind]
    ptr = (other.data_ptr() if kind == "loss" else other.A.data_ptr()) \
        if kind in ("loss", "dot", "maxabs_diff") else None
    _lib.call(name, s.pref, k, s.A.data_ptr(), ptr, out.data_ptr(),
              torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("variant", ["fused", "fused-fp16", "fused-aa-bf16"])
@pytest.mark.parametrize("shapes", [((2048, 2048), (300, 40)), ((96, 64, 80), (20, 12, 10))])
def test_noalloc_reduce_equals_reduce(variant, shapes):
    out = torch.empty(2, dtype=torch.float64, device="cuda")
    for call, shape in enumerate(shapes):  # the second call is on the smaller problem
        a, b, target = _states(variant, shape, seed=call)
        for kind in KINDS:
            other = target if kind == "loss" else b
            out.fill_(np.nan)
            _call("pylbm_hip_reduce", a, kind, other, out[0:1])
            _call("pylbm_hip_reduce_noalloc", a, kind, other, out[1:2])
            torch.cuda.synchronize()
            ref, got = out.tolist()
            print(f"{variant} {shape} {kind}: reduce {ref!r} noalloc {got!r}")
            assert np.isfinite(ref)
            assert got.hex() == ref.hex(), (variant, shape, kind)


def test_noalloc_reduce_rejects_bad_calls():
    a, b, _ = _states("fused", (64, 32), seed=3)
    out = torch.empty(1, dtype=torch.float64, device="cuda")
    lib = _lib.load()
    st = torch.cuda.current_stream().cuda_stream
    assert lib.pylbm_hip_reduce_noalloc(a.pref, 9, a.A.data_ptr(), None, out.data_ptr(),
                                        st) == 1001
    assert lib.pylbm_hip_reduce_noalloc(a.pref, _lib.RED_DOT, a.A.data_ptr(), None,
                                        out.data_ptr(), st) == 1001
    assert "second state" in lib.pylbm_hip_last_error().decode()
    # a valid call after the rejected ones still works
    _call("pylbm_hip_reduce_noalloc", a, "mass", None, out)
    assert abs(float(out.item()) - a.reduce("mass")) == 0.0
