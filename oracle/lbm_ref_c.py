"""ctypes wrapper of the C oracle (oracle/lbm_ref_c.c) -- TEST INFRASTRUCTURE ONLY.

``build()`` compiles liblbm_ref_c.so with gcc (OpenMP); ``run(case, steps, f0)`` advances a
canonical float64 state of an :class:`oracle.lbm_ref.Case`.
"""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

from . import lbm_ref as R

HERE = Path(__file__).resolve().parent
SRC = HERE / "lbm_ref_c.c"
LIB = HERE / "liblbm_ref_c.so"


class _Case(ctypes.Structure):
    _fields_ = [
        ("lattice", ctypes.c_int), ("collision", ctypes.c_int),
        ("periodic", ctypes.c_int * 3), ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("nz", ctypes.c_int), ("ulid", ctypes.c_double), ("force", ctypes.c_double * 3),
        ("force_field", ctypes.c_void_p), ("rates", ctypes.c_double * 9),
    ]


def build() -> Path:
    cmd = ["gcc", "-O2", "-fopenmp", "-shared", "-fPIC", "-o", str(LIB), str(SRC)]
    subprocess.run(cmd, check=True)
    return LIB


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = ctypes.CDLL(str(LIB))
        _lib.oracle_run.argtypes = [ctypes.POINTER(_Case), ctypes.c_int64, ctypes.c_void_p]
        _lib.oracle_run.restype = ctypes.c_int
    return _lib


def run(case: R.Case, steps: int, f0: np.ndarray | None = None) -> np.ndarray:
    f = np.ascontiguousarray(R.initial_populations(case) if f0 is None else f0,
                             dtype=np.float64).copy()
    cs = _Case()
    cs.lattice = 0 if case.dim == 2 else 1
    cs.collision = {"bgk": 0, "trt": 1, "mrt": 2}[case.collision]
    shp = tuple(case.shape) + (1,) * (3 - case.dim)
    per = tuple(case.periodic) + (True,) * (3 - case.dim)
    cs.nx, cs.ny, cs.nz = shp
    for a in range(3):
        cs.periodic[a] = int(per[a])
    cs.ulid = case.u_lid
    force = tuple(case.force) + (0.0,) * (3 - case.dim)
    for a in range(3):
        cs.force[a] = force[a]
    ff = None
    if case.force_field is not None:
        ff = np.ascontiguousarray(case.force_field, dtype=np.float64)
        cs.force_field = ff.ctypes.data
    for k, r in enumerate(case.rates()):
        cs.rates[k] = r
    rc = _load().oracle_run(ctypes.byref(cs), int(steps), f.ctypes.data)
    if rc:
        raise MemoryError("oracle_run failed")
    return f
