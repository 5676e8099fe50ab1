"""The C restatement of the oracle equals the NumPy oracle (CPU)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import lbm_ref as R
from oracle import lbm_ref_c as RC
from tests.cases import EQUIVALENCE, golden


@pytest.mark.parametrize("case", list(EQUIVALENCE))
def test_c_oracle_matches_reference_trajectories(case):
    f = RC.run(R.Case(**EQUIVALENCE[case]), 100)
    assert np.abs(f - golden("equivalence.npz")[case]).max() < 1e-13


def test_c_oracle_mrt_sq_and_force_field():
    rng = np.random.default_rng(2)
    ff = 1e-5 * rng.standard_normal((2, 10, 12))
    for kw in (dict(shape=(12, 10), tau=0.6, collision="mrt", s_e=1.2, s_eps=1.2, s_q=1.2,
                    periodic=(True, True), force=(1e-5, 2e-6)),
               dict(shape=(12, 10), tau=0.8, collision="trt", periodic=(True, False),
                    force_field=ff),
               dict(shape=(6, 5, 7), tau=0.7, collision="trt", u_lid=0.05,
                    periodic=(True, False, False))):
        c = R.Case(**kw)
        assert np.abs(RC.run(c, 40) - R.run(c, 40)).max() < 1e-13
