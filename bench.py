"""Throughput of the fused stream-collide step on one MI355X (prints one JSON line).

    python bench.py [--steps 200] [--warmup 20] [--quick] [--profile] [--no-cpu]

Workload of the headline number: D2Q9, fully periodic 4096 x 4096, fp32 populations, BGK
tau = 0.6, two-buffer fused pull kernel (hip/fused), initial state the equilibrium of the
shear wave u_x = 0.01 sin(2 pi y / 4096), u_y = 0, rho = 1 + 1e-3 U(-1, 1) per cell from a
seeded generator.  Populations are resident in HBM before timing; after a load phase of
``--warm-time`` seconds and ``warmup`` untimed steps, exactly ``steps`` launches are timed
with HIP events on the launch stream (median of three such windows).  The same line carries
every other variant of that configuration (TRT, MRT, AA, fp16/bf16 shifted storage, two
steps per launch), one step against two per launch on the walled 4096^2 cavity and channel,
the D3Q19 512^3 TRT Taylor-Green runs (fp32, fp16, bf16; pull and AA), the geometry kernels,
the forward+adjoint rollouts, the measured copy ceiling, the roofline of the dominant
kernel, the north-star ratios of this run and the float64 NumPy oracle timed on the host
(CPU baseline).  Every leg other than the headline follows the reference's protocol
(pylbm/benchmark.py:159-180): a load phase, then ``--repeats`` windows of one calibrated
step count, each at least ``--min-time`` seconds, median reported, with the GPU's clocks,
power, temperatures and limiter counters recorded per leg.

``--profile [d2q9|d2q9-2step|d2q9-2step-fp16|cavity|cavity-fp16|cavity-2step|cavity-2step-fp16|
d3q19|d3q19-fp16|d3q19-aa|d3q19-aa-fp16|geometry-d2q9|geometry-d3q19]`` runs only one step
kernel (the headline by default) and the copy kernel (for rocprofv3); the ``cavity`` forms run
the lid-driven cavity at the headline size. ``--profile sweep`` runs the sweep-stream probe
(the two-step kernel's access pattern); ``--profile copy`` runs every probe of the copy
ceiling over 4 GiB buffers.
"""

from __future__ import annotations

import argparse
import datetime
import json
import os
import platform
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "MLUPS + % HBM roofline (D2Q9 4096², D3Q19 512³); fwd+adjoint MLUPS"
SEED = 20261203


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def shear_wave_init(n: int, seed: int):
    rng = np.random.default_rng(seed)
    rho = 1.0 + 1e-3 * rng.uniform(-1.0, 1.0, size=(n, n))
    u = np.zeros((2, n, n))
    u[0] = 0.01 * np.sin(2 * np.pi * np.arange(n) / n)[:, None]
    return rho, u


def tgv_init_device(n: int, amp: float = 0.02):
    import torch

    k = torch.arange(n, dtype=torch.float64, device="cuda") * (2 * np.pi / n)
    s, c = torch.sin(k), torch.cos(k)
    u = torch.zeros((3, n, n, n), dtype=torch.float64, device="cuda")  # (a, z, y, x)
    u[0] = amp * s[None, None, :] * c[None, :, None] * c[:, None, None]
    u[1] = -amp * c[None, None, :] * s[None, :, None] * c[:, None, None]
    return None, u


def problem_2d(coll: str, n: int, rho, u):
    from pylbm_amd.problem import Problem

    extra = {"s_e": 1.2, "s_eps": 1.2} if coll == "mrt" else {}
    return Problem((n, n), tau=0.6, collision=coll, periodic=(True, True), **extra)


def run_variant(variant: str, problem, init, proto: dict, steps: int | None = None,
                warmup: int | None = None):
    """One leg: make the solver, load its initial state, then either the leg protocol
    (``time_leg``: a load phase, then ``proto['repeats']`` windows of at least
    ``proto['min_time']`` seconds) or, for the headline, exactly ``steps`` timed launches after
    ``warmup`` untimed ones, preceded by the same load phase."""
    import torch

    from pylbm_amd import backends as B
    from pylbm_amd.benchmark import gpu_state, state_delta, time_leg, time_steps, warm_to_load

    s = B.make("hip", variant, problem, init="none")
    s.init_equilibrium(*init)
    if steps is None:
        t = time_leg(s, proto["min_time"], proto["repeats"], proto["warm_time"])
    else:
        warm = warm_to_load(s, proto["warm_time"])
        s0 = gpu_state()
        t = time_steps(s, steps, warmup, repeats=3)
        t["warm_steps"] = warm
        t["machine"] = state_delta(s0, gpu_state())
    n = problem.n_nodes
    ml = [n * t["steps"] / e / 1e6 for e in t["event_s"]]
    ev = statistics.median(t["event_s"])
    rec = {
        "variant": variant, "collision": problem.collision, "lattice": problem.lattice.name,
        "size": list(problem.shape), "storage": s.storage, "steps": t["steps"],
        "repeats": len(t["event_s"]), "warm_steps": t["warm_steps"],
        "MLUPS": statistics.median(ml), "MLUPS_min": min(ml), "MLUPS_max": max(ml),
        "MLUPS_wall": n * t["steps"] / statistics.median(t["wall_s"]) / 1e6,
        "us_per_step": ev / t["steps"] * 1e6, "bytes_per_update": s.bytes_per_node,
        "machine": t["machine"],
    }
    rec["GBs"] = rec["MLUPS"] * 1e6 * s.bytes_per_node / 1e9
    ok = np.isfinite(s.reduce("mass"))
    rec["finite"] = bool(ok)
    del s
    torch.cuda.empty_cache()
    return rec


def cylinder_array_channel(n: int, coll: str = "trt", d: int = 64, pitch: int = 256):
    """D2Q9 n x n channel: y walls, parabolic velocity inlet (u_max 0.05) at x = 0, pressure
    outlet at x = n - 1, a staggered array of solid cylinders (diameter d, pitch pitch)."""
    from pylbm_amd.problem import Problem

    y, x = np.mgrid[0:n, 0:n]
    row = y // pitch  # cylinder row k: centre y = k pitch + pitch / 2
    cy = row * pitch + pitch // 2
    off = (pitch // 2) * (row % 2)  # staggered rows; centres x = off + j pitch, j >= 1
    cx = off + np.rint((x - off) / pitch).astype(np.int64) * pitch
    valid = (cx >= pitch + off) & (cx < n - pitch)
    solid = valid & ((x - cx) ** 2 + (y - cy) ** 2 <= (d / 2) ** 2)
    yy = np.arange(n) + 0.5
    inlet = 4.0 * 0.05 * yy * (n - yy) / n ** 2
    return Problem((n, n), tau=0.6, collision=coll, periodic=(False, False), solid=solid,
                   inlet=inlet, outlet=True)


def sphere_array_box(n: int, d: int = 32, pitch: int = 64):
    """D3Q19 periodic n^3 box driven by a uniform force through a cubic array of solid
    spheres (diameter d, pitch pitch)."""
    from pylbm_amd.problem import Problem

    c = (np.arange(n) % pitch) - pitch / 2 + 0.5
    r2 = c[:, None, None] ** 2 + c[None, :, None] ** 2 + c[None, None, :] ** 2
    solid = r2 <= (d / 2) ** 2
    return Problem((n, n, n), tau=0.6, collision="trt", periodic=(True, True, True),
                   force=(1e-6, 0.0, 0.0), solid=solid)


def geometry_legs(n2: int, proto: dict) -> list:
    """Obstacles + inlet + outlet (k_step_geo): the cylinder-array channel at the headline
    size and a sphere-array box in 3D; bytes per update include the link mask."""
    import torch

    out = []
    p2 = cylinder_array_channel(n2)
    for v in ("gather", "fused-fp16", "fused-bf16"):
        r = run_variant(v, p2, (None, None), proto)
        r["workload"] = f"cylinder-array channel {n2}^2 (inlet, outlet, " \
                        f"{100 * p2.solid.mean():.1f}% solid)"
        r["bytes_per_update_with_mask"] = r["bytes_per_update"] + 2
        out.append(r)
        log(f"  geometry D2Q9 {n2}^2 trt hip/{v}: {r['MLUPS']:.0f} MLUPS {r['GBs']:.0f} GB/s")
    del p2
    n3 = 256
    p3 = sphere_array_box(n3)
    for v in ("gather", "fused-fp16"):
        r = run_variant(v, p3, (None, None), proto)
        r["workload"] = f"sphere-array box {n3}^3 ({100 * p3.solid.mean():.1f}% solid)"
        r["bytes_per_update_with_mask"] = r["bytes_per_update"] + 4
        out.append(r)
        log(f"  geometry D3Q19 {n3}^3 trt hip/{v}: {r['MLUPS']:.0f} MLUPS {r['GBs']:.0f} GB/s")
    torch.cuda.empty_cache()
    return out


def walled_legs(n2: int, proto: dict) -> list:
    """One step against two steps per launch on walled D2Q9 boxes at the headline size: the
    lid-driven cavity (bounce-back on all four walls, Ladd's lid at u = 0.1) and a channel
    (periodic in x, bounce-back walls in y, uniform force 1e-6 in +x), BGK tau 0.6, fp32
    and fp16 storage, from rest."""
    from pylbm_amd.problem import Problem

    out = []
    cases = (("cavity", Problem((n2, n2), tau=0.6, collision="bgk", u_lid=0.1)),
             ("channel", Problem((n2, n2), tau=0.6, collision="bgk", periodic=(True, False),
                                 force=(1e-6, 0.0))))
    for name, p in cases:
        for v in ("fused", "fused-2step", "fused-fp16", "fused-2step-fp16"):
            r = run_variant(v, p, (None, None), proto)
            r["workload"] = f"walled {name} {n2}^2"
            out.append(r)
            log(f"  walled {name} {n2}^2 bgk hip/{v}: {r['MLUPS']:.0f} MLUPS "
                f"{r['GBs']:.0f} GB/s")
    return out


ADJOINT_CASES = ("channel", "box3d", "box3d-lam16")


def adjoint_leg(name: str, repeats: int = 3) -> dict:
    """Forward+adjoint throughput of a checkpointed differentiable rollout vs forward only.

    channel: D2Q9 BGK 1024x256 (periodic x, walls in y), tau 0.8, per-cell Guo force
             (1e-6 in +x plus 1e-7 N(0,1) noise), T = 2000, checkpoints every 100 steps,
             loss 0.5 sum |u - u_target|^2 with a smooth 4-mode target, gradients w.r.t.
             the force field and tau; states stored as shifted fp32.
    box3d:   D3Q19 BGK periodic 128^3, tau 0.7, divergence-free initial velocity (8 modes,
             amplitude 0.02), T = 500, checkpoints every 50, fp16-shifted forward storage,
             fp32 adjoint, gradient w.r.t. the initial velocity.
    box3d-lam16: box3d with the adjoint states stored as fp16 between steps (fp32
             arithmetic), which moves 114 instead of 190 bytes per node per adjoint step.
    MLUPS(fwd+adj) = nodes * T / (forward-with-checkpoints + backward time).
    """
    import torch

    from pylbm_amd import backends as B
    from pylbm_amd.adjoint import RolloutConfig, rollout, velocity_loss
    from pylbm_amd.benchmark import warm_to_load
    from pylbm_amd.fields import divergence_free_field, smooth_random_field
    from pylbm_amd.problem import Problem

    rng = np.random.default_rng(SEED)
    lam_storage = "f16" if name.endswith("-lam16") else "f32"
    if name == "channel":
        nx, ny, T, C, storage, variant = 1024, 256, 2000, 100, "f32s", "fused-f32s"
        ff = np.zeros((2, ny, nx))
        ff[0] = 1e-6
        ff += 1e-7 * rng.standard_normal((2, ny, nx))
        p = Problem((nx, ny), tau=0.8, periodic=(True, False), force_field=ff)
        target = torch.as_tensor(smooth_random_field((nx, ny), 4, 0.01, SEED + 1), device="cuda")
        F = torch.as_tensor(ff, dtype=torch.float32, device="cuda")
        u0 = None
    else:
        n, T, C, storage, variant = 128, 500, 50, "f16", "fused-fp16"
        p = Problem((n, n, n), tau=0.7, periodic=(True, True, True))
        target = torch.as_tensor(divergence_free_field(n, 8, 0.02, SEED + 2), device="cuda")
        u0 = torch.as_tensor(divergence_free_field(n, 8, 0.02, SEED + 1), device="cuda")
        F = None
    N = p.n_nodes
    # forward only: the same kernels, T steps, no checkpoints
    s = B.make("hip", variant, p, init="none")
    s.init_equilibrium(None, u0)
    warm_to_load(s, 0.25)  # the forward-only rate is taken at load, as every other leg
    fwd = []
    for _ in range(repeats):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        s.step(T)
        e1.record()
        e1.synchronize()
        fwd.append(e0.elapsed_time(e1) / 1e3)
    del s
    cfg = RolloutConfig(p, T, checkpoint=C, storage=storage, lam_storage=lam_storage)
    times, L = [], None
    for r in range(repeats + 1):
        Fr = F.clone().requires_grad_(True) if F is not None else None
        tau = torch.tensor(p.tau, dtype=torch.float64, device="cuda", requires_grad=F is not None)
        ur = u0.clone().requires_grad_(True) if u0 is not None else None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L = velocity_loss(rollout(cfg, Fr, tau, ur), target)
        L.backward()
        torch.cuda.synchronize()
        if r:  # the first pass warms up allocations
            times.append(time.perf_counter() - t0)
    torch.cuda.empty_cache()
    t_fwd, t_fa = statistics.median(fwd), statistics.median(times)
    return {"case": name, "nodes": N, "steps": T, "checkpoint": C, "storage": storage,
            "lam_storage": lam_storage,
            "forward_MLUPS": N * T / t_fwd / 1e6, "fwd_adjoint_MLUPS": N * T / t_fa / 1e6,
            "ratio": t_fwd / t_fa, "loss": float(L.detach())}


def cpu_baseline(n: int, steps: int) -> dict:
    """The float64 NumPy roll oracle on the host, D2Q9 periodic shear wave n^2."""
    from oracle import lbm_ref as R

    rho, u = shear_wave_init(n, SEED)
    case = R.Case(shape=(n, n), tau=0.6, periodic=(True, True), init_rho=rho, init_u=u)
    f = R.initial_populations(case)
    f = R.step(f, case)  # warm-up
    t0 = time.perf_counter()
    for _ in range(steps):
        f = R.step(f, case)
    dt = time.perf_counter() - t0
    return {"value": n * n * steps / dt / 1e6, "unit": "MLUPS", "cores": 1, "kind": "port",
            "sample": f"D2Q9 {n}x{n} periodic shear wave, BGK tau=0.6, {steps} steps after 1 "
                      f"warm-up step, float64 NumPy roll oracle (oracle/lbm_ref.py), "
                      f"single-threaded; host has {os.cpu_count()} logical CPUs "
                      f"({platform.processor() or platform.machine()})",
            "seconds": dt}


def north_star_ratios(variants: list, adjoint: list, head: dict, copy_gbs: float,
                      planes: dict | None = None) -> dict:
    """The north-star targets computed from this run's own legs: fp32 fraction of the copy
    ceiling, 16-bit over fp32 MLUPS per collision and pattern (legs of the same protocol),
    forward+adjoint over forward-only, each two-buffer leg's bandwidth as a fraction of the
    plane stream of its own shape (``planes``, from plane_stream_bandwidth), and the
    two-steps-per-launch legs over the one-step legs and as a fraction of the sweep stream
    (``planes``, from sweep_stream_bandwidth)."""
    def leg(variant, coll, lattice="D2Q9"):
        for r in variants:
            if (r["variant"] == variant and r["collision"] == coll and r["lattice"] == lattice
                    and r.get("protocol") != "headline" and "workload" not in r):
                return r["MLUPS"]
        return None

    def ratio(a, b):
        return None if a is None or b is None else a / b

    out = {"d2q9_fp32_frac_of_copy": head["GBs"] / copy_gbs,
           "fp16_over_fp32_target": 1.8, "fwd_adjoint_target": 0.25}
    for st in ("fp16", "bf16"):
        for coll in ("bgk", "trt", "mrt"):
            out[f"d2q9_{st}_over_fp32_{coll}"] = ratio(leg(f"fused-{st}", coll),
                                                       leg("fused", coll))
            out[f"d2q9_aa_{st}_over_fp32_{coll}"] = ratio(leg(f"fused-aa-{st}", coll),
                                                          leg("fused-aa", coll))
        out[f"d3q19_{st}_over_fp32_trt"] = ratio(leg(f"fused-{st}", "trt", "D3Q19"),
                                                 leg("fused", "trt", "D3Q19"))
        out[f"d3q19_aa_{st}_over_fp32_trt"] = ratio(leg(f"fused-aa-{st}", "trt", "D3Q19"),
                                                    leg("fused-aa", "trt", "D3Q19"))
    for r in adjoint:
        out[f"fwd_adjoint_ratio_{r['case']}"] = r["ratio"]
    if planes:
        def gbs(variant, coll, lattice):
            for r in variants:
                if (r["variant"] == variant and r["collision"] == coll
                        and r["lattice"] == lattice and "workload" not in r
                        and r.get("protocol") != "headline"):
                    return r["GBs"]
            return None

        for key, shape, g in (
                ("d2q9_fp32_bgk", "d2q9_f32", head["GBs"]),
                ("d2q9_fp16_bgk", "d2q9_16", gbs("fused-fp16", "bgk", "D2Q9")),
                ("d3q19_fp32_trt", "d3q19_f32", gbs("fused", "trt", "D3Q19")),
                ("d3q19_fp16_trt", "d3q19_16", gbs("fused-fp16", "trt", "D3Q19")),
                ("d3q19_bf16_trt", "d3q19_16", gbs("fused-bf16", "trt", "D3Q19"))):
            ref = planes.get(f"{shape}_GBs")
            out[f"{key}_frac_of_plane_stream"] = ratio(g, ref)
        # two steps per launch: the state bytes of a launch over the sweep stream's rate
        for st, shape in (("", "d2q9_sweep_f32"), ("-fp16", "d2q9_sweep_16"),
                          ("-bf16", "d2q9_sweep_16")):
            for coll in ("bgk", "trt", "mrt"):
                key = f"d2q9_2step{st.replace('-', '_')}_{coll}_frac_of_sweep_stream"
                out[key] = ratio(gbs(f"fused-2step{st}", coll, "D2Q9"),
                                 planes.get(f"{shape}_GBs"))
    for coll in ("bgk", "trt", "mrt"):
        for st in ("", "-fp16", "-bf16"):
            out[f"d2q9_2step{st.replace('-', '_')}_over_1step_{coll}"] = ratio(
                leg(f"fused-2step{st}", coll), leg(f"fused{st}", coll))

    def walled(name, variant):
        for r in variants:
            if r.get("workload", "").startswith(f"walled {name} ") and r["variant"] == variant:
                return r["MLUPS"]
        return None

    for name in ("cavity", "channel"):
        for st in ("", "-fp16"):
            out[f"walled_{name}_2step{st.replace('-', '_')}_over_1step_bgk"] = ratio(
                walled(name, f"fused-2step{st}"), walled(name, f"fused{st}"))
    return out


def load_traffic() -> dict | None:
    p = ROOT / "profiles" / "traffic_d2q9_4096_f32_bgk.json"
    if p.exists():
        return json.loads(p.read_text())
    return None


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--quick", action="store_true", help="headline + copy ceiling only")
    ap.add_argument("--profile", nargs="?", const="d2q9", default=None,
                    choices=["d2q9", "d2q9-2step", "d2q9-2step-fp16", "cavity", "cavity-fp16",
                             "cavity-2step", "cavity-2step-fp16", "d3q19", "d3q19-fp16",
                             "d3q19-aa", "d3q19-aa-fp16",
                             "geometry-d2q9", "geometry-d3q19", "copy", "planes", "sweep"],
                    help="only one step kernel and the copy kernel, for rocprofv3: the headline "
                         "(d2q9, default), its two-steps-per-launch form (fp32 / fp16), D3Q19 "
                         "512^3 TRT fp32 / fp16, two-buffer or AA, or "
                         "the geometry kernel (fp16) on the cylinder-array channel 4096^2 / "
                         "the sphere-array box 256^3")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the forward+adjoint legs")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the forward+adjoint legs (prints them, no JSON line)")
    ap.add_argument("--cache-sweep", action="store_true",
                    help="only the stream bandwidth against buffer size, 16 MiB - 4 GiB "
                         "(prints a table, no JSON line)")
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--min-time", type=float, default=0.1,
                    help="every non-headline leg: seconds per timed window (default 0.1)")
    ap.add_argument("--repeats", type=int, default=5,
                    help="every non-headline leg: timed windows, median reported (default 5)")
    ap.add_argument("--warm-time", type=float, default=0.25,
                    help="seconds of load before any leg's timing (default 0.25)")
    ap.add_argument("--n2", type=int, default=4096)
    ap.add_argument("--n3", type=int, default=512)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported by this build "
              "(multi-GPU slab decomposition is design-only).", file=sys.stderr)
        return 2

    import torch

    from pylbm_amd import _lib
    from pylbm_amd.benchmark import (SPEC_BANDWIDTH_GBS, copy_bandwidth, device_name,
                                     gpu_state, plane_stream_bandwidth,
                                     sweep_stream_bandwidth)

    if not torch.cuda.is_available():
        print("bench.py needs a HIP GPU", file=sys.stderr)
        return 1
    _lib.load()
    dev = device_name()
    log(f"device {dev}")
    if a.cache_sweep:
        from pylbm_amd.benchmark import cache_sweep

        log("stream bandwidth against buffer size (GB/s; copy = read + write)")
        for r in cache_sweep():
            log(f"{r['mib']:5d} MiB  copy nt {r['copy_nt_GBs']:6.0f}  copy {r['copy_GBs']:6.0f}  "
                f"read {r['read_GBs']:6.0f}")
        return 0
    if a.adjoint_only:
        for case in ADJOINT_CASES:
            r = adjoint_leg(case)
            log(f"fwd+adjoint {case}: {r['fwd_adjoint_MLUPS']:.0f} MLUPS vs forward "
                f"{r['forward_MLUPS']:.0f} ({100 * r['ratio']:.1f}%)")
        return 0
    n = a.n2
    if a.profile == "copy":
        # the ceiling's probe family on 4 GiB buffers, a.steps launches of each
        nbytes = 4 << 30
        x = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        y = torch.empty_like(x)
        st = int(torch.cuda.current_stream().cuda_stream)
        for _ in range(a.steps):
            _lib.call("pylbm_hip_copy", x.data_ptr(), y.data_ptr(), nbytes, st)
        for kind in range(4):
            for unroll in (1, 2, 4, 8):
                for _ in range(a.steps):
                    _lib.call("pylbm_hip_stream_probe", kind, unroll, x.data_ptr(),
                              y.data_ptr(), nbytes, st)
        torch.cuda.synchronize()
        log(f"profile: {a.steps} launches of each stream probe over {nbytes} B done")
        return 0
    if a.profile == "planes":
        # the plane-stream probe at the state shapes of D2Q9 n2^2 and D3Q19 n3^3 (fp32 and
        # 16-bit), a.steps launches each, in that order
        st = int(torch.cuda.current_stream().cuda_stream)
        for planes, eb, n_el in ((9, 4, a.n2 ** 2), (9, 2, a.n2 ** 2), (19, 4, a.n3 ** 3),
                                 (19, 2, a.n3 ** 3)):
            x = torch.empty(planes * n_el * eb // 4, dtype=torch.float32,
                            device="cuda").fill_(1.0)
            y = torch.empty_like(x)
            for _ in range(a.steps):
                _lib.call("pylbm_hip_plane_probe", planes, eb, x.data_ptr(), y.data_ptr(),
                          n_el * eb, st)
            torch.cuda.synchronize()
            del x, y
            torch.cuda.empty_cache()
            log(f"profile: {a.steps} plane probes, {planes} planes x {n_el} x {eb} B done")
        return 0
    rho, u = shear_wave_init(n, SEED)

    if a.profile:
        from pylbm_amd import backends as B
        from pylbm_amd.problem import Problem

        if a.profile == "sweep":
            # the sweep stream (fp32, n x n) a.steps times, for its kernel trace and counters
            nbytes = 9 * n * n * 4
            x = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
            y = torch.empty_like(x)
            st = int(torch.cuda.current_stream().cuda_stream)
            for _ in range(a.steps):
                _lib.call("pylbm_hip_sweep_probe", 4, x.data_ptr(), y.data_ptr(), n, n, st)
            torch.cuda.synchronize()
            log(f"profile: {a.steps} sweep streams of {n}^2 fp32 done")
            return 0
        if a.profile.startswith("d2q9"):
            p = problem_2d("bgk", n, rho, u)
            v = {"d2q9": "fused", "d2q9-2step": "fused-2step",
                 "d2q9-2step-fp16": "fused-2step-fp16"}[a.profile]
            s = B.make("hip", v, p, init="none")
            s.init_equilibrium(rho, u)
            nbytes = 9 * n * n * 4  # the bytes one step reads (and writes)
        elif a.profile.startswith("cavity"):
            # the lid-driven cavity n^2 (walls on both axes, lid 0.1), BGK tau 0.6, from rest
            p = Problem((n, n), tau=0.6, collision="bgk", u_lid=0.1)
            v = "fused" + a.profile[len("cavity"):]
            s = B.make("hip", v, p, init="none")
            s.init_equilibrium(None, None)
            nbytes = 9 * n * n * 4
        elif a.profile.startswith("geometry"):
            p = cylinder_array_channel(n) if a.profile.endswith("d2q9") else sphere_array_box(256)
            s = B.make("hip", "fused-fp16", p, init="none")
            s.init_equilibrium(None, None)
            nbytes = 1 << 30
        else:
            n3 = a.n3
            p = Problem((n3,) * 3, tau=0.55, collision="trt", periodic=(True,) * 3)
            v = {"d3q19": "fused", "d3q19-fp16": "fused-fp16", "d3q19-aa": "fused-aa",
                 "d3q19-aa-fp16": "fused-aa-fp16"}[a.profile]
            s = B.make("hip", v, p, init="none")
            s.init_equilibrium(*tgv_init_device(n3))
            nbytes = 1 << 30
        del rho, u
        s.step(a.steps)
        s.synchronize()
        log(f"profile: {a.steps} steps of hip/{s.variant} {p.lattice.name} "
            f"{'x'.join(map(str, p.shape))} done")
        x = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        y = torch.empty_like(x)
        st = int(torch.cuda.current_stream().cuda_stream)
        for _ in range(a.steps):
            _lib.call("pylbm_hip_copy", x.data_ptr(), y.data_ptr(), nbytes, st)
        torch.cuda.synchronize()
        log(f"profile: {a.steps} copies of {nbytes} B done")
        return 0

    log("copy ceiling (4 GiB buffers)")
    ceil = copy_bandwidth(4096 if not a.quick else 1024)
    cbw = ceil["copy_bandwidth_GBs"]
    log(f"copy ceiling {cbw:.0f} GB/s (best: {ceil['copy_best']}; library copy "
        f"{ceil['hip_copy_GBs']:.0f}, DtoD {ceil['memcpy_dtod_GBs']:.0f}; read-only "
        f"{ceil['read_only_GBs']:.0f}, write-only {ceil['write_only_GBs']:.0f})")
    # the same copy at the headline's own size (9 fp32 planes of n^2 each way), reported
    # beside the 4 GiB ceiling: the copy's bandwidth depends on its size
    same = copy_bandwidth(max(1, 9 * a.n2 * a.n2 * 4 // 2**20), repeats=20)
    log(f"copy at the headline's size ({same['mib']} MiB each way): "
        f"{same['copy_bandwidth_GBs']:.0f} GB/s ({same['copy_best']})")
    # the S-plane stream at the state shapes of the legs: the ceiling of their access
    # pattern (S planes read, S written per element), beside the two-stream copy
    planes = None
    if not a.quick:
        n3 = a.n3
        planes = plane_stream_bandwidth(
            {"d2q9_f32": (9, 4, n * n), "d2q9_16": (9, 2, n * n),
             "d3q19_f32": (19, 4, n3 ** 3), "d3q19_16": (19, 2, n3 ** 3)})
        # the sweep stream: the ceiling of the two-steps-per-launch kernel's access pattern
        planes.update(sweep_stream_bandwidth(n))
        log("plane streams: " + ", ".join(f"{k[:-4]} {v:.0f} GB/s" for k, v in planes.items()
                                          if k.endswith("_GBs") and "_max" not in k))

    proto = {"min_time": a.min_time, "repeats": a.repeats, "warm_time": a.warm_time}
    head = run_variant("fused", problem_2d("bgk", n, rho, u), (rho, u), proto,
                       steps=a.steps, warmup=a.warmup)
    head["protocol"] = "headline"
    log(f"headline hip/fused BGK {n}^2: {head['MLUPS']:.0f} MLUPS, {head['GBs']:.0f} GB/s "
        f"({head['steps']} timed steps)")
    variants = [head]
    if not a.quick:
        for coll in ("bgk", "trt", "mrt"):
            for v in ("fused", "fused-f32s", "fused-aa", "fused-fp16", "fused-bf16",
                      "fused-aa-fp16", "fused-aa-bf16", "fused-2step", "fused-2step-fp16",
                      "fused-2step-bf16"):
                if coll != "bgk" and v == "fused-f32s":
                    continue
                r = run_variant(v, problem_2d(coll, n, rho, u), (rho, u), proto)
                variants.append(r)
                log(f"  {coll} hip/{v}: {r['MLUPS']:.0f} MLUPS {r['GBs']:.0f} GB/s")
        from pylbm_amd.problem import Problem

        n3 = a.n3
        p3 = Problem((n3, n3, n3), tau=0.55, collision="trt", periodic=(True, True, True))
        init3 = tgv_init_device(n3)
        for v in ("fused", "fused-fp16", "fused-bf16", "fused-aa", "fused-aa-fp16",
                  "fused-aa-bf16"):
            r = run_variant(v, p3, init3, proto)
            variants.append(r)
            log(f"  D3Q19 {n3}^3 trt hip/{v}: {r['MLUPS']:.0f} MLUPS {r['GBs']:.0f} GB/s")
        del init3
        torch.cuda.empty_cache()
        variants += walled_legs(n, proto)
        variants += geometry_legs(n, proto)
    for r in variants:
        r["pct_copy"] = 100.0 * r["GBs"] / cbw
        r["pct_spec"] = 100.0 * r["GBs"] / SPEC_BANDWIDTH_GBS

    adjoint = []
    if not a.quick and not a.no_adjoint:
        for case in ADJOINT_CASES:
            r = adjoint_leg(case)
            adjoint.append(r)
            log(f"  fwd+adjoint {case}: {r['fwd_adjoint_MLUPS']:.0f} MLUPS vs forward "
                f"{r['forward_MLUPS']:.0f} ({100 * r['ratio']:.1f}%)")

    targets = north_star_ratios(variants, adjoint, head, cbw, planes)

    t_launch = head["us_per_step"] * 1e-6
    achieved = head["bytes_per_update"] * n * n / t_launch / 1e9
    traffic = load_traffic()
    roof = {"bound": "hbm", "achieved": achieved, "peak": SPEC_BANDWIDTH_GBS, "unit": "GB/s",
            "frac": achieved / SPEC_BANDWIDTH_GBS,
            "traffic": traffic["bytes_per_launch"] if traffic else None,
            "kernel": "k_step<D2Q9, BGK, f32, two-buffer, no force, single descriptor, "
                      "1 node per lane, non-temporal stores>",
            "algorithmic_bytes_per_launch": head["bytes_per_update"] * n * n,
            "avg_launch_us": t_launch * 1e6,
            "frac_of_measured_copy": achieved / cbw,
            "copy_same_size_GBs": same["copy_bandwidth_GBs"],
            "frac_of_copy_same_size": achieved / same["copy_bandwidth_GBs"]}

    cpu = None
    if not a.no_cpu and not a.quick:
        log(f"CPU baseline: NumPy float64 oracle, D2Q9 {n}^2, {a.cpu_steps} steps")
        cpu = cpu_baseline(n, a.cpu_steps)
        log(f"CPU baseline {cpu['value']:.2f} MLUPS ({cpu['seconds']:.1f} s)")

    line = {
        "metric": METRIC, "value": head["MLUPS"], "unit": "MLUPS", "n_gpus": 1,
        "steps": head["steps"], "warmup": a.warmup, "ms_per_step": head["us_per_step"] / 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": f"synthetic (seeded shear wave, seed {SEED})",
        "config": {"workload": f"D2Q9 fused pull (hip/fused), periodic {n}x{n} shear wave, "
                               "fp32 populations, BGK tau=0.6",
                   "lattice": "D2Q9", "size": [n, n], "collision": "bgk", "tau": 0.6,
                   "storage": "f32", "pattern": "ab", "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "copy_ceiling": ceil,
        "plane_stream": planes,
        "pct_copy": head["pct_copy"],
        "targets": targets,
        "protocol": {"headline": f"{a.warm_time} s of load, {a.warmup} untimed steps, then "
                                 f"3 windows of exactly {head['steps']} steps (median)",
                     "legs": f"{a.warm_time} s of load, then {a.repeats} windows of one "
                             f"calibrated step count, each >= {a.min_time} s (median; "
                             "pylbm/benchmark.py:159-180)"},
        "machine": gpu_state(),
        "date": datetime.datetime.now().astimezone().isoformat(timespec="seconds"),
        "variants": variants,
        "fwd_adjoint": adjoint,
        "device": dev,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
