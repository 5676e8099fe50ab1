"""ctypes binding of libpylbm_hip.so (the C ABI declared in include/pylbm_hip.h).

The library is loaded from the package directory, where ``pylbm_amd.build`` puts it.  It is
never replaced by a Python or CPU implementation: if it cannot be loaded, every entry point
raises and the ``hip`` backend reports itself unavailable.
"""

from __future__ import annotations

import ctypes
from pathlib import Path
from typing import Any

LIB_PATH = Path(__file__).resolve().parent / "libpylbm_hip.so"
ABI_VERSION = 4

D2Q9, D3Q19 = 0, 1
BGK, TRT, MRT = 0, 1, 2
F32, F16, BF16, F32S = 0, 1, 2, 3
AB, AA_ODD, AA_EVEN, AB2 = 0, 1, 2, 3
RED_MASS, RED_KINETIC, RED_MAXABS_DIFF, RED_LOSS, RED_DOT = 0, 1, 2, 3, 4

LATTICE_IDS = {"D2Q9": D2Q9, "D3Q19": D3Q19}
COLLISION_IDS = {"bgk": BGK, "trt": TRT, "mrt": MRT}
STORAGE_IDS = {"f32": F32, "f16": F16, "bf16": BF16, "f32s": F32S}
PATTERN_IDS = {"ab": AB, "aa-odd": AA_ODD, "aa-even": AA_EVEN, "ab2": AB2}

#: every symbol include/pylbm_hip.h declares
EXPORTS = (
    "pylbm_hip_abi_version", "pylbm_hip_last_error", "pylbm_hip_step", "pylbm_hip_run",
    "pylbm_hip_encode", "pylbm_hip_decode", "pylbm_hip_init_equilibrium",
    "pylbm_hip_macroscopic", "pylbm_hip_reduce", "pylbm_hip_copy",
    "pylbm_hip_adjoint_workspace", "pylbm_hip_adjoint_seed", "pylbm_hip_adjoint_step",
    "pylbm_hip_adjoint_init", "pylbm_hip_run_store", "pylbm_hip_adjoint_run",
    "pylbm_hip_links", "pylbm_hip_reduce_noalloc", "pylbm_hip_stream_probe",
    "pylbm_hip_plane_probe", "pylbm_hip_run_record", "pylbm_hip_adjoint_step_record",
    "pylbm_hip_adjoint_run_record", "pylbm_hip_link_groups", "pylbm_hip_run_pairs",
    "pylbm_hip_sweep_probe",
)


class Params(ctypes.Structure):
    """struct pylbm_params."""

    _fields_ = [
        ("lattice", ctypes.c_int32),
        ("collision", ctypes.c_int32),
        ("storage", ctypes.c_int32),
        ("aa", ctypes.c_int32),
        ("periodic", ctypes.c_int32 * 3),
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32),
        ("n_nodes", ctypes.c_int64),
        ("ulid", ctypes.c_float),
        ("force", ctypes.c_float * 3),
        ("rates", ctypes.c_float * 4),
        ("force_field", ctypes.c_void_p),
        ("links", ctypes.c_void_p),
        ("inlet", ctypes.c_void_p),
        ("outlet", ctypes.c_int32),
        ("lam_storage", ctypes.c_int32),
        ("link_groups", ctypes.c_void_p),
    ]


class HipError(RuntimeError):
    """A nonzero return code of the C ABI."""

    def __init__(self, code: int, msg: str) -> None:
        super().__init__(f"[{code}] {msg}")
        self.code = code


_LIB: Any = None
_LOAD_ERROR = ""


def load() -> Any:
    """Load the library once (raises OSError / RuntimeError if it is missing or stale)."""
    global _LIB, _LOAD_ERROR
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        _LOAD_ERROR = f"{LIB_PATH.name} not built (run python -m pylbm_amd.build)"
        raise OSError(_LOAD_ERROR)
    # torch owns the device memory the kernels work on: load its HIP runtime first so that
    # the library's libamdhip64 dependency binds to the runtime already in the process (a
    # second runtime initialised after torch's finds no device)
    try:
        import torch  # noqa: F401
    except ImportError:  # pragma: no cover - a C-only process uses the system runtime
        pass
    lib = ctypes.CDLL(str(LIB_PATH))
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
    pp = ctypes.POINTER(Params)
    sig = {
        "pylbm_hip_abi_version": ([], i32),
        "pylbm_hip_last_error": ([], ctypes.c_char_p),
        "pylbm_hip_step": ([pp, vp, vp, i32, vp], i32),
        "pylbm_hip_run": ([pp, vp, vp, i64, vp], i32),
        "pylbm_hip_run_pairs": ([pp, vp, vp, i64, vp], i32),
        "pylbm_hip_encode": ([pp, vp, vp, vp], i32),
        "pylbm_hip_decode": ([pp, vp, vp, vp], i32),
        "pylbm_hip_init_equilibrium": ([pp, vp, vp, vp, vp], i32),
        "pylbm_hip_macroscopic": ([pp, vp, vp, vp, vp], i32),
        "pylbm_hip_reduce": ([pp, i32, vp, vp, vp, vp], i32),
        "pylbm_hip_reduce_noalloc": ([pp, i32, vp, vp, vp, vp], i32),
        "pylbm_hip_copy": ([vp, vp, i64, vp], i32),
        "pylbm_hip_stream_probe": ([i32, i32, vp, vp, i64, vp], i32),
        "pylbm_hip_plane_probe": ([i32, i32, vp, vp, i64, vp], i32),
        "pylbm_hip_sweep_probe": ([i32, vp, vp, i32, i32, vp], i32),
        "pylbm_hip_adjoint_workspace": ([pp], i64),
        "pylbm_hip_adjoint_seed": ([pp, vp, vp, vp, vp, vp], i32),
        "pylbm_hip_adjoint_step": ([pp, vp, vp, vp, vp, vp, vp], i32),
        "pylbm_hip_adjoint_init": ([pp, vp, vp, vp, vp, vp, vp], i32),
        "pylbm_hip_run_store": ([pp, vp, ctypes.POINTER(vp), i64, vp], i32),
        "pylbm_hip_adjoint_run": ([pp, ctypes.POINTER(vp), i64, vp, vp, vp, vp,
                                   ctypes.POINTER(ctypes.c_int), vp], i32),
        "pylbm_hip_links": ([pp, vp, vp, vp], i32),
        "pylbm_hip_link_groups": ([pp, vp, vp, vp], i32),
        "pylbm_hip_run_record": ([pp, vp, ctypes.POINTER(vp), ctypes.POINTER(vp), i64, vp],
                                 i32),
        "pylbm_hip_adjoint_step_record": ([pp, vp, vp, vp, vp, vp], i32),
        "pylbm_hip_adjoint_run_record": ([pp, ctypes.POINTER(vp), i64, vp, vp, vp,
                                          ctypes.POINTER(ctypes.c_int), vp], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    v = lib.pylbm_hip_abi_version()
    if v != ABI_VERSION:
        raise RuntimeError(f"libpylbm_hip ABI {v} != expected {ABI_VERSION}; rebuild")
    _LIB = lib
    return lib


def load_error() -> str:
    try:
        load()
        return ""
    except Exception as e:  # noqa: BLE001
        return f"{type(e).__name__}: {e}"


def check(rc: int) -> None:
    if rc != 0:
        msg = load().pylbm_hip_last_error().decode(errors="replace")
        raise HipError(rc, msg)


def call(name: str, *args: Any) -> None:
    """Call an entry point and raise HipError on a nonzero return."""
    check(getattr(load(), name)(*args))
