"""pylbm_amd: pylbm's fused lattice-Boltzmann time step as hand-written gfx950 HIP kernels.

The package mirrors the reference's host interface (``Problem``, the backend registry and
the ``Solver`` API, pylbm/problem.py, pylbm/backends/) and registers the ``hip`` backend
with variants ``fused``, ``fused-aa``, ``fused-fp16``, ``fused-bf16``, ``fused-aa-fp16``
and ``fused-aa-bf16``, plus ``fused-f32s`` (fp32 storage of f - w), ``gather`` (obstacles,
inlet, outlet) and ``fused-2step[-fp16|-bf16]`` (two steps per launch on periodic D2Q9
boxes).  Kernels live in libpylbm_hip.so (``python -m pylbm_amd.build``).
"""

__version__ = "0.1.0"
