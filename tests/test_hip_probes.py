"""The bandwidth probes move exactly the bytes they claim (GPU).

The copy ceiling (pylbm_hip_copy, pylbm_hip_stream_probe), the plane-stream ceiling
(pylbm_hip_plane_probe) and the sweep-stream ceiling (pylbm_hip_sweep_probe) are only
ceilings if every probe touches every byte (the sweep probe's destination must be one
streaming step of its source): each copy
form must leave dst equal to src, byte for byte, on a destination filled with NaN first;
the write-only stream must fill dst; the read-only stream must leave a 1.0f-filled dst
untouched beyond its one word per workgroup.  Sizes exercise partial last workgroups
(copies) and both the XCD-ordered and the plain workgroup order (plane probe: 8 * chunk
divides the workgroup count or not).
"""

from __future__ import annotations

import pytest
import torch

from pylbm_amd import _lib

pytestmark = pytest.mark.gpu


def _stream() -> int:
    return int(torch.cuda.current_stream().cuda_stream)


def _pattern(nbytes: int) -> torch.Tensor:
    g = torch.Generator(device="cpu").manual_seed(nbytes)
    return torch.randint(0, 2**31 - 1, (nbytes // 4,), generator=g, dtype=torch.int32).cuda()


@pytest.mark.parametrize("nbytes", [16, 4096 * 16 + 48, 3 << 20])
def test_library_copy_is_exact(nbytes):
    src = _pattern(nbytes)
    dst = torch.full_like(src, -1)
    _lib.call("pylbm_hip_copy", src.data_ptr(), dst.data_ptr(), nbytes, _stream())
    torch.cuda.synchronize()
    assert torch.equal(src, dst)


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("unroll", [1, 2, 4, 8])
def test_copy_probes_are_exact(kind, unroll):
    nbytes = (256 * unroll * 5 + 7) * 16  # a partial last workgroup
    src = _pattern(nbytes)
    dst = torch.full_like(src, -1)
    _lib.call("pylbm_hip_stream_probe", kind, unroll, src.data_ptr(), dst.data_ptr(), nbytes,
              _stream())
    torch.cuda.synchronize()
    assert torch.equal(src, dst)


@pytest.mark.parametrize("unroll", [1, 8])
def test_write_and_read_probes(unroll):
    nbytes = (256 * unroll * 3 + 5) * 16
    dst = torch.zeros(nbytes // 4, dtype=torch.float32, device="cuda")
    _lib.call("pylbm_hip_stream_probe", 3, unroll, 0, dst.data_ptr(), nbytes, _stream())
    torch.cuda.synchronize()
    assert bool((dst == 1.0).all())
    # the read-only stream of a 1.0f fill never stores: dst keeps its contents
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    out = torch.full(((nbytes // (256 * unroll) + 16) // 4,), 7.0, device="cuda")
    _lib.call("pylbm_hip_stream_probe", 2, unroll, src.data_ptr(), out.data_ptr(), nbytes,
              _stream())
    torch.cuda.synchronize()
    assert bool((out == 7.0).all())


@pytest.mark.parametrize("planes", [9, 19])
@pytest.mark.parametrize("elem_bytes", [2, 4])
@pytest.mark.parametrize("elems", [256 * 3, 256 * 8 * 16 * 2])
def test_plane_probe_copies_every_plane(planes, elem_bytes, elems):
    plane = elems * elem_bytes
    src = _pattern(planes * plane)
    dst = torch.full_like(src, -1)
    _lib.call("pylbm_hip_plane_probe", planes, elem_bytes, src.data_ptr(), dst.data_ptr(),
              plane, _stream())
    torch.cuda.synchronize()
    assert torch.equal(src, dst)


@pytest.mark.parametrize("args", [(8, 4, 1024), (19, 8, 1024), (9, 4, 1000), (19, 2, 510)])
def test_plane_probe_rejects_bad_shapes(args):
    planes, eb, plane = args
    buf = torch.zeros(64 * 1024, dtype=torch.int32, device="cuda")
    with pytest.raises(_lib.HipError):
        _lib.call("pylbm_hip_plane_probe", planes, eb, buf.data_ptr(), buf.data_ptr(), plane,
                  _stream())


# D2Q9 velocities in the library's order (rest, E, N, W, S, NE, NW, SW, SE)
_C9 = [(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1)]


@pytest.mark.parametrize("elem_bytes", [4, 2])
@pytest.mark.parametrize("shape", [(600, 71), (256, 10), (1, 3), (4096, 37)])
def test_sweep_probe_is_one_streaming_step(elem_bytes, shape):
    """The sweep-stream probe reads and writes every element it is timed for: dst is one
    periodic pull-streaming step of src, dst[i][y][x] = src[i][y - c_y][x - c_x], bit for
    bit, on a NaN-filled destination (partial last tiles, chunks of rows with a shorter
    last one, a one-column state)."""
    nx, ny = shape
    n = nx * ny
    dt = torch.int32 if elem_bytes == 4 else torch.int16
    g = torch.Generator(device="cpu").manual_seed(nx * 1000 + ny)
    src = torch.randint(-2**15, 2**15 - 1, (9, ny, nx), generator=g).to(dt).cuda()
    dst = torch.full_like(src, -1)
    _lib.call("pylbm_hip_sweep_probe", elem_bytes, src.data_ptr(), dst.data_ptr(), nx, ny,
              _stream())
    torch.cuda.synchronize()
    want = torch.stack([torch.roll(src[i], shifts=(cy, cx), dims=(0, 1))
                        for i, (cx, cy) in enumerate(_C9)])
    assert torch.equal(dst, want), f"{shape} {elem_bytes} B, {n} nodes"
