"""Measurement helpers: copy-bandwidth ceiling, event-timed step loops, traffic model.

Follows the reference's timing rules (pylbm/benchmark.py:1-7,159-236): warm-up before
timing, device synchronisation around every timed region, and the minimal traffic model of
2 q s bytes per node update (pylbm/backends/base.py:90-92).  Timing uses HIP events
recorded on torch's current stream, the stream every kernel of this package is launched on.
"""

from __future__ import annotations

import statistics
import time
from typing import Any

from . import _lib

SPEC_BANDWIDTH_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def _torch():
    import torch

    return torch


def copy_bandwidth(mib: int = 4096, repeats: int = 20) -> dict:
    """Median device copy bandwidth (read + write counted, bench/roofline.py:30-51).

    Two probes over ``mib`` MiB buffers (default 4 GiB, far larger than the 256 MB
    Infinity Cache): the library's streaming copy kernel and torch ``copy_``
    (hipMemcpyDtoD).  Returns GB/s for both and the per-launch kernel time.
    """
    torch = _torch()
    nbytes = mib * 2**20
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    stream = int(torch.cuda.current_stream().cuda_stream)

    def hip_copy() -> None:
        _lib.call("pylbm_hip_copy", a.data_ptr(), b.data_ptr(), nbytes, stream)

    def dtod() -> None:
        b.copy_(a)

    out: dict[str, Any] = {"mib": mib}
    for name, fn in (("hip_copy", hip_copy), ("memcpy_dtod", dtod)):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        times = []
        for _ in range(repeats):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            times.append(e0.elapsed_time(e1) / 1e3)
        bw = [2 * nbytes / t / 1e9 for t in times]
        out[f"{name}_GBs"] = statistics.median(bw)
        out[f"{name}_max_GBs"] = max(bw)
        out[f"{name}_ms"] = statistics.median(times) * 1e3
    del a, b
    torch.cuda.empty_cache()
    out["copy_bandwidth_GBs"] = max(out["hip_copy_GBs"], out["memcpy_dtod_GBs"])
    return out


def time_steps(solver: Any, steps: int, warmup: int, repeats: int = 1) -> dict:
    """Event-timed ``solver.step(steps)`` after ``warmup`` untimed steps.

    Returns per-repeat seconds; each repeat is bracketed by synchronisations and HIP events
    on the current stream, so the event time is the device time of exactly ``steps``
    kernel launches (one launch per step; AA alternates odd/even launches).
    """
    torch = _torch()
    m = solver.step_multiple
    warmup -= warmup % m
    steps -= steps % m
    if warmup:
        solver.step(warmup)
    solver.synchronize()
    ev, wall = [], []
    for _ in range(repeats):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0.record()
        solver.step(steps)
        e1.record()
        torch.cuda.synchronize()
        wall.append(time.perf_counter() - t0)
        ev.append(e0.elapsed_time(e1) / 1e3)
    return {"steps": steps, "event_s": ev, "wall_s": wall}


def mlups(n_nodes: int, steps: int, seconds: float) -> float:
    return n_nodes * steps / seconds / 1e6
