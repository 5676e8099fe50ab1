"""Measurement helpers: copy-bandwidth ceiling, event-timed step loops, traffic model.

Follows the reference's timing rules (pylbm/benchmark.py:1-7,159-236): warm-up before
timing, device synchronisation around every timed region, and the minimal traffic model of
2 q s bytes per node update (pylbm/backends/base.py:90-92).  Timing uses HIP events
recorded on torch's current stream, the stream every kernel of this package is launched on.
"""

from __future__ import annotations

import math
import statistics
import time
from typing import Any

from . import _lib

SPEC_BANDWIDTH_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def _torch():
    import torch

    return torch


#: pylbm_hip_stream_probe kinds
PROBE_KINDS = {"copy_nt": 0, "copy": 1, "read": 2, "write": 3}


def _event_times(fn, repeats: int) -> list[float]:
    torch = _torch()
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
    return times


def _load(fn, seconds: float) -> None:
    """Run ``fn`` back to back for ``seconds`` of wall time: the clocks leave the idle
    power state before anything is timed."""
    torch = _torch()
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(8):
            fn()
        torch.cuda.synchronize()


#: plane-stream shapes: name -> (planes, element bytes, elements per plane), the state
#: arrays of the benchmark's workloads (D2Q9 4096^2 and D3Q19 512^3, fp32 and 16-bit)
PLANE_SHAPES = {"d2q9_f32": (9, 4, 4096 * 4096), "d2q9_16": (9, 2, 4096 * 4096),
                "d3q19_f32": (19, 4, 512 ** 3), "d3q19_16": (19, 2, 512 ** 3)}


def plane_stream_bandwidth(shapes: dict | None = None, repeats: int = 20,
                           load_s: float = 0.1) -> dict:
    """Rate of the S-plane stream (``pylbm_hip_plane_probe``) at each shape: S planes read
    and S written per element, counted as 2 S element bytes, the bytes a two-buffer step of
    that lattice moves.  The ceiling for a step kernel of that shape: the probe has the
    kernel's memory access pattern and no arithmetic."""
    torch = _torch()
    out: dict[str, Any] = {}
    stream = int(torch.cuda.current_stream().cuda_stream)
    for name, (planes, eb, n) in (shapes or PLANE_SHAPES).items():
        plane = n * eb
        a = torch.empty(planes * plane // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)

        def fn(a=a, b=b, planes=planes, eb=eb, plane=plane):
            _lib.call("pylbm_hip_plane_probe", planes, eb, a.data_ptr(), b.data_ptr(), plane,
                      stream)

        _load(fn, load_s)
        times = _event_times(fn, repeats)
        bw = [2 * planes * plane / t / 1e9 for t in times]
        out[f"{name}_GBs"] = statistics.median(bw)
        out[f"{name}_max_GBs"] = max(bw)
        out[f"{name}_us"] = statistics.median(times) * 1e6
        del a, b
        torch.cuda.empty_cache()
    return out


def sweep_stream_bandwidth(n: int = 4096, repeats: int = 20, load_s: float = 0.1) -> dict:
    """Rate of the sweep stream (``pylbm_hip_sweep_probe``) over a periodic n x n D2Q9 state
    of fp32 and of 16-bit elements: the memory access pattern of the two-steps-per-launch
    kernel without its arithmetic, counted as 2 x 9 n^2 element bytes per launch (the state
    read and written once).  The ceiling for that kernel, beside the plane stream of the
    one-step kernel (``plane_stream_bandwidth``)."""
    torch = _torch()
    out: dict[str, Any] = {}
    stream = int(torch.cuda.current_stream().cuda_stream)
    for name, eb in (("d2q9_sweep_f32", 4), ("d2q9_sweep_16", 2)):
        nbytes = 9 * n * n * eb
        a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)

        def fn(a=a, b=b, eb=eb):
            _lib.call("pylbm_hip_sweep_probe", eb, a.data_ptr(), b.data_ptr(), n, n, stream)

        _load(fn, load_s)
        times = _event_times(fn, repeats)
        bw = [2 * nbytes / t / 1e9 for t in times]
        out[f"{name}_GBs"] = statistics.median(bw)
        out[f"{name}_max_GBs"] = max(bw)
        out[f"{name}_us"] = statistics.median(times) * 1e6
        del a, b
        torch.cuda.empty_cache()
    return out


def copy_bandwidth(mib: int = 4096, repeats: int = 20, sweep: bool = True) -> dict:
    """Median device copy bandwidth (read + write counted, bench/roofline.py:30-51).

    Probes over ``mib`` MiB buffers (default 4 GiB, far larger than the 256 MB Infinity
    Cache): the library's streaming copy (``pylbm_hip_copy``), torch ``copy_``
    (hipMemcpyDtoD) and, with ``sweep``, the probe family of ``pylbm_hip_stream_probe``:
    copy with non-temporal or ordinary accesses at 1, 2, 4 and 8 16-byte accesses per lane,
    and read-only / write-only streams (their bytes counted once).  The ceiling is the best
    copy of them all; the read-only and write-only rates are reported beside it.
    """
    torch = _torch()
    nbytes = mib * 2**20
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    stream = int(torch.cuda.current_stream().cuda_stream)
    probes: dict[str, tuple[Any, int]] = {
        "hip_copy": (lambda: _lib.call("pylbm_hip_copy", a.data_ptr(), b.data_ptr(), nbytes,
                                       stream), 2),
        "memcpy_dtod": (lambda: b.copy_(a), 2),
    }
    if sweep:
        for name, kind in PROBE_KINDS.items():
            for u in (1, 2, 4, 8):
                probes[f"{name}_u{u}"] = (
                    (lambda k=kind, u=u: _lib.call("pylbm_hip_stream_probe", k, u,
                                                   a.data_ptr(), b.data_ptr(), nbytes,
                                                   stream)),
                    2 if name.startswith("copy") else 1)
    out: dict[str, Any] = {"mib": mib}
    _load(probes["hip_copy"][0], 0.25)
    for name, (fn, ways) in probes.items():
        times = _event_times(fn, repeats)
        bw = [ways * nbytes / t / 1e9 for t in times]
        out[f"{name}_GBs"] = statistics.median(bw)
        out[f"{name}_max_GBs"] = max(bw)
        out[f"{name}_ms"] = statistics.median(times) * 1e3
    del a, b
    torch.cuda.empty_cache()
    copies = [k for k in probes if probes[k][1] == 2]
    best = max(copies, key=lambda k: out[f"{k}_GBs"])
    out["copy_bandwidth_GBs"] = out[f"{best}_GBs"]
    out["copy_best"] = best
    if sweep:
        out["read_only_GBs"] = max(out[f"read_u{u}_GBs"] for u in (1, 2, 4, 8))
        out["write_only_GBs"] = max(out[f"write_u{u}_GBs"] for u in (1, 2, 4, 8))
    return out


def cache_sweep(sizes_mib: tuple = (16, 32, 64, 96, 128, 192, 256, 512, 4096),
                repeats: int = 20) -> list[dict]:
    """Stream bandwidth against buffer size: the copy (non-temporal and ordinary accesses)
    and the read-only stream of ``pylbm_hip_stream_probe`` at one 16-byte access per lane,
    over buffers from well inside the 256 MB Infinity Cache to far beyond it (median of
    ``repeats`` launches after a 0.05 s load phase per probe; a copy counts read + write)."""
    torch = _torch()
    stream = int(torch.cuda.current_stream().cuda_stream)
    rows = []
    for mib in sizes_mib:
        nbytes = mib << 20
        a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        # read-only stream: at most one 16-byte word per workgroup lands in dst
        r = torch.empty(nbytes // 1024 + 64, dtype=torch.float32, device="cuda")
        row: dict[str, Any] = {"mib": mib}
        for name, kind, dst, ways in (("copy_nt", 0, b, 2), ("copy", 1, b, 2),
                                      ("read", 2, r, 1)):
            def fn(kind=kind, dst=dst):
                _lib.call("pylbm_hip_stream_probe", kind, 1, a.data_ptr(), dst.data_ptr(),
                          nbytes, stream)
            _load(fn, 0.05)
            t = statistics.median(_event_times(fn, repeats))
            row[f"{name}_GBs"] = ways * nbytes / t / 1e9
            row[f"{name}_us"] = t * 1e6
        rows.append(row)
        del a, b, r
        torch.cuda.empty_cache()
    return rows


_SMI: dict[str, Any] = {}


def gpu_state() -> dict:
    """Read-only machine snapshot of GPU 0 (pylbm/benchmark.py:121-130 reads the same
    quantities from nvidia-smi): graphics and memory clocks (MHz), socket power (W), hotspot
    and HBM temperatures (C), and the firmware's accumulated power- and thermal-limit
    counters (their difference over a timed window says whether a limiter was active).

    Uses the ``amdsmi`` Python module when it is importable; returns ``{}`` otherwise or on
    any error, so a missing tool never stops a measurement.  The firmware refreshes these
    values every few tens of milliseconds, so they describe a leg, not a launch.
    """
    try:
        import amdsmi

        if "h" not in _SMI:
            amdsmi.amdsmi_init()
            _SMI["h"] = amdsmi.amdsmi_get_processor_handles()[0]
        h = _SMI["h"]
        m = amdsmi.amdsmi_get_gpu_metrics_info(h)
        out = {"gfxclk_MHz": m.get("current_gfxclk"), "uclk_MHz": m.get("current_uclk"),
               "socket_power_W": m.get("current_socket_power"),
               "hotspot_C": m.get("temperature_hotspot"), "hbm_C": m.get("temperature_mem")}
        try:
            v = amdsmi.amdsmi_get_violation_status(h)
            for k in ("acc_counter", "acc_ppt_pwr", "acc_socket_thrm", "acc_hbm_thrm",
                      "acc_vr_thrm", "acc_prochot_thrm"):
                if isinstance(v.get(k), int):
                    out[k] = v[k]
        except Exception:  # noqa: BLE001 - older firmware has no violation table
            pass
        return {k: val for k, val in out.items() if isinstance(val, (int, float))}
    except Exception:  # noqa: BLE001 - a snapshot must never fail a measurement
        return {}


def device_name() -> str:
    """Product name of GPU 0: the ASIC's market name from ``amdsmi`` when it reads one
    (torch reports the generic "AMD Radeon Graphics" on some drivers), else torch's name."""
    try:
        import amdsmi

        if "h" not in _SMI:
            amdsmi.amdsmi_init()
            _SMI["h"] = amdsmi.amdsmi_get_processor_handles()[0]
        name = str(amdsmi.amdsmi_get_gpu_asic_info(_SMI["h"]).get("market_name", "")).strip()
        if name and name.upper() != "N/A":
            return name
    except Exception:  # noqa: BLE001 - fall back to torch's name
        pass
    return _torch().cuda.get_device_name(0)


def state_delta(s0: dict, s1: dict) -> dict:
    """Snapshot after a window, plus the growth of each accumulated limiter counter."""
    out = dict(s1)
    for k in s1:
        if k.startswith("acc_") and k in s0:
            out[k] = s1[k] - s0[k]
    return out


def warm_to_load(solver: Any, seconds: float) -> int:
    """Keep the GPU busy with the solver's own kernel for at least ``seconds`` of wall time.

    On MI355X the 16-bit step kernels run up to 30% slower for about 30 ms after the GPU
    leaves idle (a power-management transient of the firmware: the kernel's duration on one
    fixed input rises from 102 to 133 us and recovers, while 150 ms of any prior load
    removes it; fp32 kernels are not affected).  A leg therefore starts its timed windows
    only after a load phase measured in time, not in steps.  Returns the steps taken.
    """
    m = solver.step_multiple
    chunk, done = 2 * m, 0
    t0 = time.perf_counter()
    while True:
        solver.step(chunk)
        solver.synchronize()
        done += chunk
        dt = time.perf_counter() - t0
        if dt >= seconds:
            return done
        want = (seconds - dt) * done / max(dt, 1e-6)  # steps that fill the rest at this rate
        chunk = max(m, min(8 * chunk, math.ceil(want / m) * m))


def time_leg(solver: Any, min_time: float = 0.1, repeats: int = 5,
             warm_time: float = 0.25) -> dict:
    """One benchmark leg: a load phase of ``warm_time`` seconds, then ``repeats`` timed
    windows of the same step count, each at least ``min_time`` long (the protocol of
    pylbm/benchmark.py:159-180 with HIP events on the launch stream).

    Returns the per-window event and wall times, the steps per window and the machine
    state (clocks, power, temperatures, limiter counters over the windows).
    """
    torch = _torch()
    m = solver.step_multiple
    warm_steps = warm_to_load(solver, warm_time)
    # calibrate: the warm phase's rate sizes a window of >= min_time with 10% margin
    probe = max(m, 2 * m)
    t0 = time.perf_counter()
    solver.step(probe)
    solver.synchronize()
    dt = max(time.perf_counter() - t0, 1e-6)
    steps = max(probe, math.ceil(1.1 * min_time / dt * probe / m) * m)
    s0 = gpu_state()
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
    return {"steps": steps, "event_s": ev, "wall_s": wall, "warm_steps": warm_steps,
            "machine": state_delta(s0, gpu_state())}


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


# ----------------------------------------------------------------------------- plans
# The reference's benchmark engine (pylbm/benchmark.py:67-90, 154-236, 242-324): a
# lid-driven cavity (Re = 1000, U = 0.05, all walls) per size, warm-up, >= 5 repeats of
# >= min_time seconds, median / min / max MLUPS, effective bandwidth from the 2 q s model.
PLANS = {
    "quick": dict(sizes_2d=[1024], sizes_3d=[64], repeats=3, min_time=0.15, copy_mib=1024),
    "dev": dict(sizes_2d=[256, 1024, 4096], sizes_3d=[64, 128, 256], repeats=5, min_time=0.5,
                copy_mib=4096),
    "full": dict(sizes_2d=[256, 512, 1024, 2048, 4096], sizes_3d=[64, 128, 256, 512],
                 repeats=5, min_time=0.5, copy_mib=4096),
}
VARIANTS = ["fused", "fused-f32s", "fused-aa", "fused-fp16", "fused-bf16", "fused-aa-fp16",
            "fused-aa-bf16"]


def bench_problem(dim: int, n: int, collision: str = "bgk"):
    from .problem import Problem

    return Problem.cavity(n, re=1000.0, u_lid=0.05, collision=collision, dim=dim)


def _wall(solver: Any, steps: int) -> float:
    t0 = time.perf_counter()
    solver.step(steps)
    solver.synchronize()
    return time.perf_counter() - t0


def time_solver(solver: Any, min_time: float, repeats: int) -> tuple[list[float], int]:
    """Per-repeat wall times and steps per repeat (the protocol of pylbm/benchmark.py:159-180:
    a warm-up, repeats of at least ``min_time`` seconds, synchronised on both sides).

    The repeat length comes from probes whose step count grows eightfold until one takes a
    tenth of ``min_time``; the rate it measures then sizes the repeat with 20% margin.
    """
    m = solver.step_multiple
    warm_to_load(solver, 0.25)  # warm-up: leaves the GPU at load (see warm_to_load)
    probe = 2 * m
    while True:
        dt = _wall(solver, probe)
        if dt >= 0.1 * min_time:
            break
        probe *= 8
    steps = max(probe, math.ceil(1.2 * min_time / dt * probe / m) * m)
    return [_wall(solver, steps) for _ in range(repeats)], steps


def measure(variant: str, dim: int, n: int, plan: dict, copy_bw: float,
            collision: str = "bgk", log: Any = print) -> dict | None:
    """One (variant, size) record in the reference's schema (pylbm/benchmark.py:183-236)."""
    from . import backends as B

    torch = _torch()
    cls = B.get("hip", variant)
    ok, why = cls.available()
    if not ok:
        log(f"  skip hip/{variant}: {why}")
        return None
    p = bench_problem(dim, n, collision)
    ok, why = cls.supports(p)
    if not ok:
        log(f"  skip hip/{variant} {n}^{dim}: {why}")
        return None
    s = cls(p)
    state0 = gpu_state()
    times, steps = time_solver(s, plan["min_time"], plan["repeats"])
    ml = [p.n_nodes * steps / t / 1e6 for t in times]
    med = statistics.median(ml)
    eff = med * 1e6 * s.bytes_per_node / 1e9
    rec = {"backend": "hip", "variant": variant, "lattice": p.lattice.name,
           "collision": collision, "size": list(p.shape), "dtype": "float32",
           "storage": s.storage, "metric": "MLUPS", "value": med, "repeats": len(times),
           "median": med, "min": min(ml), "max": max(ml), "steps_per_repeat": steps,
           "bytes_per_unit": s.bytes_per_node, "effective_GBs": eff,
           "pct_copy_bw": 100.0 * eff / copy_bw, "pct_spec_bw": 100.0 * eff / SPEC_BANDWIDTH_GBS,
           "flops_per_unit": None, "pct_peak_flops": None,
           "memory_MiB": s.memory_bytes() / 2**20,
           "notes": "; ".join([cls.description,
                               f"gfx clock {state0.get('gfxclk_MHz', '?')} MHz, "
                               f"hotspot {state0.get('hotspot_C', '?')} C"])}
    del s
    torch.cuda.empty_cache()
    log(f"  hip {variant:14s} {collision} {'x'.join(map(str, p.shape)):>14s} {med:10.1f} MLUPS "
        f"{eff:7.1f} GB/s {rec['pct_copy_bw']:5.1f}% copy")
    return rec


def run(plan_name: str = "quick", out: str | None = None, log: Any = print) -> dict:
    """Run a benchmark plan over every hip variant; write JSON (pylbm/benchmark.py:242-324)."""
    import datetime
    import json
    import platform

    torch = _torch()
    plan = PLANS[plan_name]
    t0 = time.perf_counter()
    ceil = copy_bandwidth(plan["copy_mib"])
    cbw = ceil["copy_bandwidth_GBs"]
    log(f"[pylbm_amd bench:{plan_name}] measured copy bandwidth {cbw:.1f} GB/s")
    runs = []
    for dim, sizes in ((2, plan["sizes_2d"]), (3, plan["sizes_3d"])):
        for v in VARIANTS:
            for n in sizes:
                try:
                    r = measure(v, dim, n, plan, cbw, log=log)
                except Exception as e:  # noqa: BLE001 - one failure never ends the run
                    log(f"  FAIL hip/{v} {n}^{dim}: {type(e).__name__}: {e}")
                    continue
                if r:
                    runs.append(r)
    for coll in ("trt", "mrt"):
        n = plan["sizes_2d"][-1]
        r = measure("fused", 2, n, plan, cbw, collision=coll, log=log)
        if r:
            runs.append(r)
    date = datetime.datetime.now().astimezone().isoformat(timespec="seconds")
    result = {"machine": {"python": platform.python_version(), "date": date,
                          "torch": torch.__version__, "gpu": device_name()},
              "ceiling": {"copy_bandwidth_GBs": cbw, "spec_bandwidth_GBs": SPEC_BANDWIDTH_GBS,
                          "fp32_tflops": None, **ceil},
              "plan": plan_name, "provisional": False,
              "wall_time_s": time.perf_counter() - t0, "runs": runs}
    if out:
        with open(out, "w") as fh:
            json.dump(result, fh, indent=1)
        log(f"[pylbm_amd bench] wrote {out} ({len(runs)} runs)")
    return result
