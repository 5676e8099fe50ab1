/*
 * pylbm_hip.h -- C ABI of libpylbm_hip.so, the gfx950 (MI355X) fused lattice-Boltzmann
 * stream-collide library behind pylbm's `hip` backend.
 *
 * Conventions (identical to the reference so the canonical state is interchangeable):
 *   - populations are SoA planes f[i * N + n], n = x + nx * (y + ny * z), x contiguous;
 *     N = nx * ny * nz; plane offsets are 64-bit, in-plane indices 32-bit (N < 2^31);
 *   - D2Q9 order: 0 rest, E, N, W, S, NE, NW, SW, SE; D3Q19 opposite pairs adjacent
 *     (pylbm/lattice.py:49-67);
 *   - storage PYLBM_F32 keeps f; PYLBM_F16 / PYLBM_BF16 keep h = f - w_i (arithmetic fp32);
 *     PYLBM_F32S keeps h = f - w_i in fp32 (no reference counterpart: the precision the
 *     differentiable rollouts use for the non-equilibrium part);
 *   - AA layout (params.aa = 1): the canonical post-collision population i sits in slot
 *     opp(i) (pylbm/backends/_fused.py:83-103);
 *   - the moving lid is the top wall of the last axis, moving in +x (Ladd term).
 *
 * Every entry point returns 0 on success or a nonzero code (a hipError_t, or one of the
 * PYLBM_E* codes below); pylbm_hip_last_error() returns the message of the last failure
 * on the calling thread.  `stream` is a hipStream_t passed as void* (NULL = null stream).
 * All pointers are device pointers unless stated otherwise.  Launches are asynchronous.
 */
#ifndef PYLBM_HIP_H
#define PYLBM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PYLBM_HIP_ABI_VERSION 4

enum { PYLBM_D2Q9 = 0, PYLBM_D3Q19 = 1 };
enum { PYLBM_BGK = 0, PYLBM_TRT = 1, PYLBM_MRT = 2 };
enum { PYLBM_F32 = 0, PYLBM_F16 = 1, PYLBM_BF16 = 2, PYLBM_F32S = 3 };
enum { PYLBM_AB = 0, PYLBM_AA_ODD = 1, PYLBM_AA_EVEN = 2, PYLBM_AB2 = 3 };
enum {
  PYLBM_RED_MASS = 0, PYLBM_RED_KINETIC = 1, PYLBM_RED_MAXABS_DIFF = 2, PYLBM_RED_LOSS = 3,
  PYLBM_RED_DOT = 4
};

enum {
  PYLBM_EINVAL = 1001,   /* bad parameters (message says which) */
  PYLBM_ERANGE = 1002,   /* size outside what the kernels index */
  PYLBM_ENOTSUP = 1003   /* combination not implemented (e.g. MRT on D3Q19) */
};

/*
 * Kernel parameters of one box problem.  Mirrors the compile-time KernelSpec and runtime
 * scalars of the reference (pylbm/codegen.py:36-51, pylbm/backends/_fused.py:129-136,
 * pylbm/backends/cupy_backend.py:51-70); the runtime rates follow runtime_rates
 * (pylbm/codegen.py:437-445): BGK r0 = 1/tau; TRT (r0, r1) = (s+, s-);
 * MRT (r0, r1, r2, r3) = (s_nu, s_e, s_eps, s_q).
 */
typedef struct pylbm_params {
  int32_t lattice;      /* PYLBM_D2Q9 | PYLBM_D3Q19 */
  int32_t collision;    /* PYLBM_BGK | PYLBM_TRT | PYLBM_MRT (D2Q9 only) */
  int32_t storage;      /* PYLBM_F32 | PYLBM_F16 | PYLBM_BF16 | PYLBM_F32S */
  int32_t aa;           /* 1: AA layout (slot opp(i) holds population i), else 0 */
  int32_t periodic[3];  /* per axis x, y, z: 1 periodic, 0 half-way bounce-back walls */
  int32_t nx, ny, nz;   /* nz = 1 for D2Q9 */
  int64_t n_nodes;      /* N = nx * ny * nz (checked) */
  float ulid;           /* lid speed (top wall of the last axis), 0 = none */
  float force[3];       /* uniform Guo body force (ignored when force_field != NULL) */
  float rates[4];       /* r0..r3, see above */
  const float* force_field; /* NULL, or per-cell Guo force, SoA (d, N) fp32, device */
  /*
   * General geometry (pylbm/problem.py:37-40, semantics pylbm/core.py:56-148).  The
   * reference's fused kernels take box problems only (pylbm/problem.py:90-92); here the
   * two-buffer pattern (PYLBM_AB) also runs these, with arrays under 4 GiB in total.
   */
  const void* links;    /* NULL, or the per-node solid-source mask of pylbm_hip_links */
  const float* inlet;   /* NULL, or u_x(y) of the velocity inlet on the x = 0 wall (ny
                           fp32 values; D2Q9; no effect when x is periodic) */
  int32_t outlet;       /* 1: rho = 1 pressure outlet on the x = nx - 1 column (D2Q9) */
  int32_t lam_storage;  /* element type of the adjoint states lam of the pylbm_hip_adjoint_*
                           entries: PYLBM_F32 (0) or PYLBM_F16 (values rounded to nearest
                           fp16 on store, all arithmetic fp32); the forward entries ignore it.
                           No reference counterpart. */
  const uint32_t* link_groups; /* NULL, or the group flags of pylbm_hip_link_groups for
                           `links` (requires links): a 16-bit step then issues the bounce-back
                           loads of every group that needs them together with its pull
                           loads, and skips the mask load of the groups that do not (fp32
                           steps ignore them).  Same results either way.  No reference
                           counterpart. */
} pylbm_params;

/* ABI version (PYLBM_HIP_ABI_VERSION of the library). */
int pylbm_hip_abi_version(void);

/* Message of the last nonzero return on this thread ("" if none). Host pointer. */
const char* pylbm_hip_last_error(void);

/*
 * One fused time step: pull streaming + bounce-back + Ladd lid, moments, equilibrium,
 * Guo forcing, BGK/TRT/MRT collision, store.  pattern = PYLBM_AB reads src writes dst
 * (src != dst); PYLBM_AA_ODD / PYLBM_AA_EVEN work in place (src == dst required).
 * PYLBM_AB2 advances two steps in one launch, src -> dst, with the intermediate state on
 * chip (the same values as two PYLBM_AB steps); D2Q9 box problems (periodic axes or walls,
 * the lid; no links, inlet or outlet) whose state is under 4 GiB only (PYLBM_ENOTSUP
 * otherwise).  No reference counterpart for PYLBM_AB2.
 * Replaces FusedSolver._launch (pylbm/backends/_fused.py:108-110,
 * pylbm/backends/cupy_backend.py:101-103) and the kernel lbm_kernel
 * (pylbm/backends/cuda/lbm.cu:50-53).
 */
int pylbm_hip_step(const pylbm_params* p, const void* src, void* dst, int pattern,
                   void* stream);

/*
 * nsteps time steps enqueued from C.  Two-buffer (p->aa == 0): ping-pong a -> b -> a ...;
 * after an odd nsteps the state is in b.  AA (p->aa == 1): nsteps must be even, b is
 * ignored, pairs of odd/even steps in place in a.
 * Replaces FusedSolver._advance (pylbm/backends/_fused.py:116-124) and Solver.step
 * (pylbm/backends/base.py:68-73).
 */
int pylbm_hip_run(const pylbm_params* p, void* a, void* b, int64_t nsteps, void* stream);

/*
 * nsteps (even) time steps as nsteps / 2 launches of the PYLBM_AB2 pattern, ping-ponging
 * a -> b -> a ...; after an odd nsteps / 2 the state is in b.  D2Q9 box problems with a
 * state under 4 GiB (PYLBM_ENOTSUP otherwise).  The step loop of the
 * two-steps-per-launch variants; no reference counterpart (a form of
 * FusedSolver._advance, pylbm/backends/_fused.py:116-124).
 */
int pylbm_hip_run_pairs(const pylbm_params* p, void* a, void* b, int64_t nsteps, void* stream);

/*
 * nsteps two-buffer steps from src, writing the state after step k to dst[k] (a host array
 * of nsteps device pointers; each dst[k] must differ from the buffer it is computed from).
 * Enqueues the whole loop from C: used by the differentiable rollouts to store checkpoints
 * and to recompute a checkpoint segment.  No reference counterpart (the reference has no
 * rollout storage); a generalisation of FusedSolver._advance (pylbm/backends/_fused.py:116-124).
 */
int pylbm_hip_run_store(const pylbm_params* p, const void* src, void* const* dst,
                        int64_t nsteps, void* stream);

/*
 * pylbm_hip_run_store that also writes moment records: when rec[k] (a host array of nsteps
 * device pointers, or rec itself) is not NULL, step k writes to rec[k] the density and
 * velocity its collision used at every node, (rho, u_x, u_y, u_z) fp32 (u = (sum c g + F/2)
 * / rho of the pulled populations g; u_z = 0 for D2Q9), node-major: N x 16 bytes.  The
 * adjoint of a rollout that does not differentiate tau needs only these
 * (pylbm_hip_adjoint_run_record), 16 bytes per node instead of the stored state.
 * BGK / TRT box problems whose (q, N) fp32 array spans less than 4 GiB (PYLBM_ERANGE
 * otherwise).  No reference counterpart (the reference has no rollout storage); a
 * generalisation of FusedSolver._advance (pylbm/backends/_fused.py:116-124).
 */
int pylbm_hip_run_record(const pylbm_params* p, const void* src, void* const* dst,
                         void* const* rec, int64_t nsteps, void* stream);

/*
 * Canonical float64 state (q, N) -> stored layout (shift by -w_i for 16-bit storage,
 * rounded float64 -> float32 -> storage; AA slot swap).
 * Replaces FusedSolver._encode / set_populations (pylbm/backends/_fused.py:83-93).
 */
int pylbm_hip_encode(const pylbm_params* p, const double* canon, void* f, void* stream);

/*
 * Stored layout -> canonical float64 state (q, N) (exact widening, + w_i, AA unswap).
 * Replaces FusedSolver.populations (pylbm/backends/_fused.py:95-103).
 */
int pylbm_hip_decode(const pylbm_params* p, const void* f, double* canon, void* stream);

/*
 * Equilibrium initial state written in the stored layout, computed in float64 then
 * encoded exactly as pylbm_hip_encode.  rho: N values or NULL (rho = 1); u: (d, N) values
 * or NULL (u = 0).  Replaces Problem.initial_populations + _encode
 * (pylbm/problem.py:108-115, pylbm/lattice.py:112-119) without a host round trip.
 */
int pylbm_hip_init_equilibrium(const pylbm_params* p, const double* rho, const double* u,
                               void* f, void* stream);

/*
 * Post-collision observables in float64: rho (N) and u (d, N) with u = (sum c f - F/2)/rho.
 * Either output may be NULL.  Replaces Solver.macroscopic -> core.macroscopic
 * (pylbm/backends/base.py:79-83, pylbm/core.py:255-266).
 */
int pylbm_hip_macroscopic(const pylbm_params* p, const void* f, double* rho, double* u,
                          void* stream);

/*
 * Global float64 reductions of a stored state into *out (device pointer, one double):
 *   PYLBM_RED_MASS        sum over nodes and populations of f
 *   PYLBM_RED_KINETIC     sum over nodes of 0.5 rho |u|^2 (post-collision u)
 *   PYLBM_RED_MAXABS_DIFF max |f - g| over all canonical values, other = second state
 *                         with the same params
 *   PYLBM_RED_LOSS        0.5 sum over nodes of |u - u_target|^2, other = u_target as
 *                         float64 (d, N) (the rollout loss)
 *   PYLBM_RED_DOT         sum over all canonical values of f g, other = second state with
 *                         the same params (with storage PYLBM_F32, any two fp32 (q, N)
 *                         arrays, e.g. adjoint states)
 * No reference counterpart as a kernel: the reference computes these on the host from
 * populations() (pylbm/validation/taylor_green.py:43-44, tests/test_reference.py:10-15).
 */
int pylbm_hip_reduce(const pylbm_params* p, int kind, const void* f, const void* other,
                     double* out, void* stream);

/*
 * pylbm_hip_reduce without any allocation: the same kernels and the same result, bit for
 * bit, with the per-workgroup partials kept in a static device array of the library
 * (4096 doubles, 32 KiB per GPU, held until the process ends) instead of a stream-ordered
 * allocation.  It calls no hipMalloc, hipMallocAsync or hipFree and has no counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops).  Every problem size the library accepts fits (the reduction grid
 * is capped at 4096 workgroups).  All calls on one GPU share its array: calls that overlap
 * on one GPU (different streams or host threads) race; serialise them.
 */
int pylbm_hip_reduce_noalloc(const pylbm_params* p, int kind, const void* f,
                             const void* other, double* out, void* stream);

/*
 * Per-node solid-source mask of a solid field: links[n] bit i (i >= 1) is set when the
 * node x - c_i, wrapped periodically on every axis, is solid (solid: N bytes, nonzero =
 * solid; links: N uint16 for D2Q9, N uint32 for D3Q19).  Pass links in params.links to
 * step with obstacles.  Replaces the per-direction rolled solid masks of make_geometry
 * (pylbm/core.py:56-71) and the q*N int64 gather table of the torch gather solvers
 * (pylbm/backends/torch_backend.py:80-103).
 */
int pylbm_hip_links(const pylbm_params* p, const uint8_t* solid, void* links, void* stream);

/*
 * Group flags of a link mask: groups[g] (uint32) for the g-th group of 64 consecutive nodes
 * of a row, g = (y + ny * z) * ceil(nx / 64) + x / 64, is nonzero when a node of the group has
 * a solid pull source (links[n] != 0) or lies on a wall of a non-periodic x axis, else 0
 * (groups: ny * nz * ceil(nx / 64) values).  Pass them in params.link_groups (beside
 * params.links).  No reference counterpart: the reference has no fused obstacle kernel
 * (pylbm/core.py:56-78 runs obstacles in its unfused solvers).
 */
int pylbm_hip_link_groups(const pylbm_params* p, const void* links, uint32_t* groups,
                          void* stream);

/*
 * Streaming copy of nbytes (multiple of 16) from src to dst: the measured copy ceiling.
 * Replaces the torch copy_ bandwidth probe (bench/roofline.py:30-51).
 */
int pylbm_hip_copy(const void* src, void* dst, int64_t nbytes, void* stream);

/*
 * Probes of the bandwidth ceiling, nbytes a multiple of 16: kind 0 copy with non-temporal
 * loads and stores (pylbm_hip_copy is kind 0 at unroll 1), 1 copy with ordinary loads and
 * stores, 2 read-only stream of src (dst receives at most one 16-byte word per
 * workgroup, only for contents a fill of 1.0f never has: give it nbytes / (256 unroll) + 16
 * bytes), 3 write-only stream into dst (src unused); unroll = 1, 2, 4 or 8 16-byte accesses
 * per lane.  Extends the copy probe of bench/roofline.py:30-51, which has one form only.
 */
int pylbm_hip_stream_probe(int kind, int unroll, const void* src, void* dst, int64_t nbytes,
                           void* stream);

/*
 * Plane-stream probe: the access pattern of a two-buffer step without its arithmetic.
 * src and dst each hold `planes` (9 or 19) planes of plane_bytes (a multiple of 256
 * elements of elem_bytes = 2 or 4); lane x reads element x of every plane of src, then
 * stores each to the same place in dst (non-temporal), with the step kernel's workgroup
 * order.  dst becomes a copy of src.  Its rate is the ceiling of an S-plane SoA stream of
 * that shape, beside the two-stream copy of pylbm_hip_copy.  Extends the copy probe of
 * bench/roofline.py:30-51, which has one form only.
 */
int pylbm_hip_plane_probe(int planes, int elem_bytes, const void* src, void* dst,
                          int64_t plane_bytes, void* stream);

/*
 * Sweep-stream probe: the access pattern of the two-steps-per-launch kernel (PYLBM_AB2)
 * without its arithmetic.  src and dst each hold a periodic nx x ny D2Q9 state of
 * elem_bytes (2 or 4) elements, 9 planes of nx ny, under 4 GiB; workgroups of 256 columns
 * sweep chunks of rows as that kernel does, loading a row ahead, and store each row
 * pulled from src: dst becomes one periodic streaming step of src,
 * dst[i][y][x] = src[i][y - c_y,i][x - c_x,i].  Its rate (2 * 9 nx ny elem_bytes per
 * launch) is the ceiling of that kernel's memory pattern.  Extends the copy probe of
 * bench/roofline.py:30-51, which has one form only.
 */
int pylbm_hip_sweep_probe(int elem_bytes, const void* src, void* dst, int nx, int ny,
                          void* stream);

/* ---------------------------------------------------------------------------------------
 * Discrete adjoint of the two-buffer step (BGK / TRT, Guo forcing; params->aa must be 0).
 * No reference counterpart: the reference has no gradient path (its torch rungs run under
 * torch.no_grad(), pylbm/backends/torch_backend.py:74,128,152,175,257).  Used by the
 * checkpointed rollout behind torch.autograd.Function (pylbm_amd/adjoint.py):
 *   seed -> for t = T-1 .. 0: adjoint_step(f^t) -> init.
 * Adjoint states lam are (q, N) arrays of p->lam_storage elements (fp32 or fp16); dF is fp32
 * (d, N), accumulated in place.
 */

/* Doubles of the per-workgroup dL/ds+- partial-sum buffer pylbm_hip_adjoint_step needs
 * (2 per workgroup), or -1 on bad params. */
int64_t pylbm_hip_adjoint_workspace(const pylbm_params* p);

/* Terminal adjoint of a loss L(u(f_T)) given grad_u = dL/du (fp32, (d, N)), with
 * u = (sum c f - F/2)/rho the post-collision velocity: writes lam = dL/df_T and, with a
 * force field, initialises dF = dL/dF of that velocity definition (-grad_u / (2 rho)). */
int pylbm_hip_adjoint_seed(const pylbm_params* p, const void* f_T, const float* grad_u,
                           void* lam, float* dF, void* stream);

/* One adjoint step: lam_out = (d f^{t+1} / d f^t)^T lam_in, with f_t the stored state the
 * forward step read.  Accumulates dF += dL/dF(x) (force field only) and adds the step's
 * dL/ds+ and dL/ds- (BGK: both into slot 0 pattern [s+, 0]) to dtau_part[2 g], [2 g + 1]
 * for workgroup g; the caller zeroes dtau_part once and sums it at the end.  dtau_part may
 * be NULL: the rate sensitivities are then not computed (a rollout that does not
 * differentiate tau). */
int pylbm_hip_adjoint_step(const pylbm_params* p, const void* f_t, const void* lam_in,
                           void* lam_out, float* dF, double* dtau_part, void* stream);

/* nsteps adjoint steps over a stored segment, in reverse: for k = nsteps-1 .. 0,
 * lam_out = adjoint of step(states[k]) applied to lam_in, ping-ponging lam_a (input on
 * entry) and lam_b; *which = 0 if the final adjoint state is in lam_a, 1 if in lam_b.
 * Pairs of steps run in place in lam_a (an odd last step goes to lam_b); the result equals
 * the ping-pong of pylbm_hip_adjoint_step bit for bit.
 * dF and dtau_part accumulate as in pylbm_hip_adjoint_step (dtau_part NULL: none). */
int pylbm_hip_adjoint_run(const pylbm_params* p, const void* const* states, int64_t nsteps,
                          void* lam_a, void* lam_b, float* dF, double* dtau_part, int* which,
                          void* stream);

/* pylbm_hip_adjoint_step and pylbm_hip_adjoint_run with the forward quantities taken from
 * the steps' moment records (pylbm_hip_run_record) instead of the stored states: the same
 * adjoint algebra, one 16-byte load per node in place of the q-population pull.  No rate
 * sensitivities (they need the pulled populations): for rollouts that do not differentiate
 * tau.  Sizes as pylbm_hip_run_record.  On D3Q19 boxes periodic on every axis with rows of
 * at most 128 nodes, fp32 adjoint states and no force field, pylbm_hip_adjoint_run_record
 * runs two steps per launch, the first step's result held on chip, ping-ponging lam_a and
 * lam_b; the result equals the single steps bit for bit. */
int pylbm_hip_adjoint_step_record(const pylbm_params* p, const void* rec_t, const void* lam_in,
                                  void* lam_out, float* dF, void* stream);
int pylbm_hip_adjoint_run_record(const pylbm_params* p, const void* const* recs, int64_t nsteps,
                                 void* lam_a, void* lam_b, float* dF, int* which, void* stream);

/* Gradient of the initial equilibrium f0 = feq(rho0, u0) (float64 fields; NULL rho0 = 1,
 * NULL u0 = 0): du0 (d, N) and drho0 (N) in float64 from lam0 = dL/df0 (either output may
 * be NULL). */
int pylbm_hip_adjoint_init(const pylbm_params* p, const double* rho0, const double* u0,
                           const void* lam0, double* du0, double* drho0, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PYLBM_HIP_H */
