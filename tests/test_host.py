"""Host-side logic (CPU): problem description, rates, tables, registry, C-ABI surface.

No test here launches a kernel: the C-ABI checks only exercise parameter validation, which
returns before any HIP call.
"""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import lbm_ref as R
from pylbm_amd import _lib
from pylbm_amd import backends as B
from pylbm_amd.lattice import D2Q9, D3Q19, MRT_M_D2Q9, MRT_NORM_D2Q9, equilibrium, magic_rate
from pylbm_amd.problem import Problem

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "pylbm_hip.h"

# ----------------------------------------------------------------------------- tables


@pytest.mark.parametrize("mine,ref", [(D2Q9, R.D2Q9), (D3Q19, R.D3Q19)])
def test_lattice_tables_equal_oracle(mine, ref):
    assert np.array_equal(mine.c, ref.c)
    assert np.array_equal(mine.w, ref.w)
    assert np.array_equal(mine.opp, ref.opp)
    assert (mine.d, mine.q) == (ref.d, ref.q)


def test_mrt_tables_equal_oracle():
    assert np.array_equal(MRT_M_D2Q9, R.MRT_M)
    assert np.array_equal(MRT_NORM_D2Q9, R.MRT_NORM)


def test_kernel_tables_match_python_tables():
    """The HIP lattice tables (csrc/lbm_common.hpp) carry the same velocities / opposites."""
    src = (ROOT / "pylbm_amd" / "csrc" / "lbm_common.hpp").read_text()
    for lat, name in ((D2Q9, "D2Q9"), (D3Q19, "D3Q19")):
        block = src[src.index(f"struct {name}"):]
        m = re.search(r"constexpr int t\[\d+\]\[3\] = \{(.*?)\};", block, re.S)
        rows = re.findall(r"\{\s*(-?\d+),\s*(-?\d+),\s*(-?\d+)\s*\}", m.group(1))
        c = np.array(rows, dtype=int)[:, : lat.d]
        assert np.array_equal(c, lat.c)
    m = re.search(r"constexpr int t\[9\] = \{([^}]*)\};", src[src.index("struct D2Q9"):])
    assert [int(v) for v in m.group(1).split(",")] == list(D2Q9.opp)
    # D3Q19 opposite rule used by the kernel: pairs (1,2), (3,4), ...
    assert [0] + [i + 1 if i % 2 else i - 1 for i in range(1, 19)] == list(D3Q19.opp)
    mm = re.search(r"constexpr int t\[9\]\[9\] = \{(.*?)\};", src, re.S)
    vals = [int(v) for v in re.findall(r"-?\d+", mm.group(1))]
    assert np.array_equal(np.array(vals).reshape(9, 9), MRT_M_D2Q9)


def test_equilibrium_matches_oracle():
    rng = np.random.default_rng(0)
    rho = 1 + 0.01 * rng.standard_normal((3, 4, 5))
    u = 0.03 * rng.standard_normal((3, 3, 4, 5))
    assert np.array_equal(equilibrium(D3Q19, rho, u), R.equilibrium(R.D3Q19, rho, u))


# ----------------------------------------------------------------------------- problem


def test_cavity_tau_and_derived_rates():
    p = Problem.cavity(128, re=100.0, u_lid=0.1)
    assert abs(p.tau - 0.884) < 1e-12
    assert p.runtime_rates() == (1 / p.tau, 0.0, 0.0, 0.0)
    sp, sm = Problem((8, 8), tau=0.6, collision="trt").trt_rates()
    assert np.isclose(sm, 1 / 2.375) and np.isclose(sp, 1 / 0.6)
    assert np.isclose(Problem((8, 8, 8), tau=0.55, collision="trt").trt_rates()[1], 1 / 4.25)


@pytest.mark.parametrize("kw", [
    dict(shape=(8, 6), tau=0.7, collision="trt"),
    dict(shape=(8, 6), tau=0.6, collision="mrt", s_e=1.2, s_eps=1.2),
    dict(shape=(8, 6), tau=0.6, collision="mrt", s_e=1.2, s_eps=1.2, s_q=1.2),
    dict(shape=(8, 6, 4), tau=0.55, collision="trt", magic=0.25),
])
def test_runtime_rates_match_oracle_rates(kw):
    p, c = Problem(**kw), R.Case(**kw)
    r = c.rates()
    if p.collision == "trt":
        assert np.allclose(p.runtime_rates()[:2], r)
    else:
        assert np.allclose(p.runtime_rates(), (r[7], r[1], r[2], r[4]))
    assert np.isclose(magic_rate(1 / p.tau, p.magic), R.magic_rate(1 / p.tau, p.magic))


@pytest.mark.parametrize("kw,err", [
    (dict(shape=(8, 8), tau=0.5), "tau"),
    (dict(shape=(8, 8, 8), collision="mrt"), "MRT"),
    (dict(shape=(8, 8), periodic=(True, True), u_lid=0.1), "lid"),
    (dict(shape=(8,)), "2D or 3D"),
    (dict(shape=(8, 8), force=(1e-5, 0.0), force_field=np.zeros((2, 8, 8))), "either"),
    (dict(shape=(8, 8), force_field=np.zeros((2, 4, 4))), "shape"),
])
def test_problem_rejects(kw, err):
    with pytest.raises(ValueError, match=err):
        Problem(**kw)


def test_initial_populations_match_oracle():
    rng = np.random.default_rng(1)
    rho = 1 + 1e-3 * rng.uniform(-1, 1, (6, 7))
    u = 0.01 * rng.standard_normal((2, 6, 7))
    p = Problem((7, 6), init_rho=rho, init_u=u)
    c = R.Case((7, 6), init_rho=rho, init_u=u)
    assert np.array_equal(p.initial_populations(), R.initial_populations(c))


# ----------------------------------------------------------------------------- registry


def test_registry_never_crashes():
    st = B.status()
    assert len(st) == len(B.entries()) == 11
    for b, v, ok, why in st:
        assert b == "hip"
        assert ok or why  # unavailable entries say why


def test_unknown_entry_raises_keyerror():
    with pytest.raises(KeyError):
        B.get("hip", "nope")


def test_supports_geometry_on_two_buffer_variants_only():
    solid = np.zeros((8, 8), dtype=bool)
    geo = Problem((8, 8), solid=solid, inlet=np.zeros(8), outlet=True)
    for v in ("fused", "fused-f32s", "fused-fp16", "fused-bf16", "gather"):
        assert B.get("hip", v).supports(geo)[0], v
    for v in ("fused-aa", "fused-aa-fp16", "fused-aa-bf16", "fused-2step", "fused-2step-fp16",
              "fused-2step-bf16"):
        ok, why = B.get("hip", v).supports(geo)
        assert not ok and "box" in why
    cls = B.get("hip", "fused")
    assert cls.supports(Problem((8, 8)))[0]
    assert cls.supports(Problem((8, 8), collision="mrt"))[0]
    # geometry needs the whole state under 4 GiB (single-descriptor addressing)
    big = Problem((1 << 14, 1 << 13), solid=None, outlet=True)
    ok, why = cls.supports(big)
    assert not ok and "4 GiB" in why


def test_make_raises_unsupported_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(B.Unsupported):
        B.make("hip", "fused", Problem((8, 8)))


# ----------------------------------------------------------------------------- C ABI


def _declared_functions() -> list[str]:
    text = HEADER.read_text()
    return re.findall(r"^\w[\w\s\*]*?\b(pylbm_hip_\w+)\s*\(", text, re.M)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _declared_functions()
    assert sorted(declared) == sorted(_lib.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.pylbm_hip_abi_version() == _lib.ABI_VERSION


def test_params_struct_layout():
    # int32 x 10, int32 nz, int64 n_nodes at offset 40, floats, then a pointer
    assert _lib.Params.n_nodes.offset == 40
    assert _lib.Params.ulid.offset == 48
    assert _lib.Params.rates.offset == 64
    assert _lib.Params.force_field.offset == 80
    assert _lib.Params.links.offset == 88
    assert _lib.Params.inlet.offset == 96
    assert _lib.Params.outlet.offset == 104
    assert _lib.Params.lam_storage.offset == 108
    assert _lib.Params.link_groups.offset == 112
    assert ctypes.sizeof(_lib.Params) == 120


def _params(**over) -> _lib.Params:
    p = _lib.Params()
    p.lattice, p.collision, p.storage, p.aa = _lib.D2Q9, _lib.BGK, _lib.F32, 0
    p.periodic[0] = p.periodic[1] = p.periodic[2] = 1
    p.nx, p.ny, p.nz, p.n_nodes = 16, 8, 1, 128
    p.rates[0] = 1.0
    for k, v in over.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("over,needle", [
    (dict(lattice=7), "unknown lattice"),
    (dict(collision=_lib.MRT, lattice=_lib.D3Q19, nz=2, n_nodes=256), "MRT"),
    (dict(storage=9), "unknown storage"),
    (dict(n_nodes=100), "n_nodes"),
    (dict(nz=2, n_nodes=256), "nz == 1"),
    (dict(nx=0, n_nodes=0), "sizes"),
    (dict(aa=2), "aa must be"),
    (dict(lam_storage=2), "lam_storage"),
    (dict(outlet=2), "outlet must be"),
    (dict(aa=1, links=0x3000), "two-buffer"),
    (dict(lattice=_lib.D3Q19, nz=2, n_nodes=256, outlet=1), "D2Q9 only"),
    (dict(nx=1, n_nodes=8, outlet=1), "nx >= 2"),
])
def test_abi_rejects_bad_params(over, needle):
    lib = _lib.load()
    p = _params(**over)
    dummy = ctypes.c_void_p(0x1000)
    rc = lib.pylbm_hip_step(ctypes.byref(p), dummy, ctypes.c_void_p(0x2000), _lib.AB, None)
    assert rc != 0
    assert needle in lib.pylbm_hip_last_error().decode()


def test_abi_rejects_bad_calls():
    lib = _lib.load()
    p = _params()
    a, b = ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000)
    assert lib.pylbm_hip_step(ctypes.byref(p), a, a, _lib.AB, None) == 1001
    assert "src != dst" in lib.pylbm_hip_last_error().decode()
    assert lib.pylbm_hip_step(ctypes.byref(p), a, b, _lib.AA_ODD, None) == 1001
    assert lib.pylbm_hip_step(ctypes.byref(p), a, b, 5, None) == 1001
    paa = _params(aa=1)
    assert lib.pylbm_hip_run(ctypes.byref(paa), a, None, 3, None) == 1001
    assert "pairs" in lib.pylbm_hip_last_error().decode()
    assert lib.pylbm_hip_run(ctypes.byref(p), a, a, 2, None) == 1001
    assert lib.pylbm_hip_copy(a, b, 17, None) == 1001
    assert lib.pylbm_hip_reduce(ctypes.byref(p), 9, a, None, b, None) == 1001
    for kind in (_lib.RED_MAXABS_DIFF, _lib.RED_DOT, _lib.RED_LOSS):
        assert lib.pylbm_hip_reduce(ctypes.byref(p), kind, a, None, b, None) == 1001
    lid = _params(ulid=0.1)
    assert lib.pylbm_hip_step(ctypes.byref(lid), a, b, _lib.AB, None) == 1001
    assert "lid" in lib.pylbm_hip_last_error().decode()
    big = _params(nx=65536, ny=65536, n_nodes=65536 * 65536)
    assert lib.pylbm_hip_step(ctypes.byref(big), a, b, _lib.AB, None) == 1002


def test_abi_rejects_fp32_side_planes_beyond_4gib():
    """A 16-bit state of more than 2^30 nodes fits the kernels' 32-bit byte offsets, but
    its fp32 force field and fp32 adjoint planes (4 bytes per node) do not: such calls are
    rejected before any HIP call instead of wrapping the offsets."""
    lib = _lib.load()
    a, b = ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000)
    nx, ny = 65536, 16385  # N = 1,073,807,360 > 2^30, N * 2 < 2^32
    n = nx * ny
    ok16 = _params(storage=_lib.F16, nx=nx, ny=ny, n_nodes=n)
    # the 16-bit state alone passes the checks (the workspace query makes no HIP call)
    assert lib.pylbm_hip_adjoint_workspace(ctypes.byref(ok16)) > 0
    field = _params(storage=_lib.F16, nx=nx, ny=ny, n_nodes=n, force_field=0x5000)
    assert lib.pylbm_hip_step(ctypes.byref(field), a, b, _lib.AB, None) == 1002
    assert "force-field" in lib.pylbm_hip_last_error().decode()
    assert lib.pylbm_hip_run(ctypes.byref(field), a, b, 2, None) == 1002
    for lam in (_lib.F32, _lib.F16):
        adj = _params(storage=_lib.F16, nx=nx, ny=ny, n_nodes=n, lam_storage=lam)
        assert lib.pylbm_hip_adjoint_step(ctypes.byref(adj), a, b, ctypes.c_void_p(0x3000),
                                          None, ctypes.c_void_p(0x4000), None) == 1002
        assert "adjoint plane" in lib.pylbm_hip_last_error().decode()
        assert lib.pylbm_hip_adjoint_seed(ctypes.byref(adj), a, b, ctypes.c_void_p(0x3000),
                                          None, None) == 1002
    # just under the limit: 2^30 - 256 nodes with a force field pass the range checks
    under = _params(storage=_lib.F16, nx=65536, ny=16383, n_nodes=65536 * 16383,
                    force_field=0x5000)
    assert lib.pylbm_hip_adjoint_workspace(ctypes.byref(under)) > 0


def test_abi_rejects_bad_record_calls():
    """The moment-record entries check their sizes, collision and geometry before any HIP
    call: MRT and obstacles are not covered, and every array must stay under 4 GiB."""
    lib = _lib.load()
    a, b, c = ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000), ctypes.c_void_p(0x3000)
    arr = (ctypes.c_void_p * 2)(0x5000, 0x6000)
    which = ctypes.c_int(0)

    def last():
        return lib.pylbm_hip_last_error().decode()

    mrt = _params(collision=_lib.MRT)
    assert lib.pylbm_hip_run_record(ctypes.byref(mrt), a, arr, arr, 2, None) == 1003
    assert "BGK and TRT" in last()
    geo = _params(links=0x7000)
    assert lib.pylbm_hip_run_record(ctypes.byref(geo), a, arr, arr, 2, None) == 1003
    assert "box geometries" in last()
    assert lib.pylbm_hip_adjoint_step_record(ctypes.byref(geo), a, b, c, None, None) == 1003
    assert lib.pylbm_hip_run_record(ctypes.byref(_params(aa=1)), a, arr, arr, 2, None) == 1001
    # D3Q19 512 x 512 x 256 with 16-bit states: the state is under 4 GiB, a q-plane fp32
    # array (the adjoint states) is not
    big = _params(lattice=_lib.D3Q19, storage=_lib.F16, nx=512, ny=512, nz=256,
                  n_nodes=512 * 512 * 256)
    assert lib.pylbm_hip_run_record(ctypes.byref(big), a, arr, arr, 2, None) == 1002
    assert "4 GiB" in last()
    assert lib.pylbm_hip_adjoint_step_record(ctypes.byref(big), a, b, c, None, None) == 1002
    assert lib.pylbm_hip_adjoint_run_record(ctypes.byref(big), arr, 2, b, c, None,
                                            ctypes.byref(which), None) == 1002
    p = _params()
    assert lib.pylbm_hip_adjoint_step_record(ctypes.byref(p), a, b, b, None, None) == 1001
    assert lib.pylbm_hip_adjoint_run_record(ctypes.byref(p), arr, 2, b, b, None,
                                            ctypes.byref(which), None) == 1001
    assert lib.pylbm_hip_run_record(ctypes.byref(p), a, arr, arr, -1, None) == 1001


def test_abi_rejects_link_groups_without_links():
    lib = _lib.load()
    a, b = ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000)
    p = _params(link_groups=0x3000)
    assert lib.pylbm_hip_step(ctypes.byref(p), a, b, _lib.AB, None) == 1001
    assert "link_groups needs links" in lib.pylbm_hip_last_error().decode()
    assert lib.pylbm_hip_link_groups(ctypes.byref(_params()), None, b, None) == 1001
    assert lib.pylbm_hip_link_groups(ctypes.byref(_params()), a, None, None) == 1001
