// Shared device-side definitions of the gfx950 fused LBM kernels: lattice traits, storage
// codecs, the per-node collision (BGK / TRT / MRT with Guo forcing) and kernel arguments.
//
// Lattice orderings, weights and the D2Q9 MRT basis are the reference's
// (pylbm/lattice.py:49-87); the collision arithmetic restates pylbm/core.py:151-247 in
// fp32, the precision of the reference's fused kernels (pylbm/backends/cuda/lbm.cu:100-163).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pylbm_hip.h"

namespace plbm {

constexpr int kBlock = 256;  // 4 wave64 per workgroup

// ------------------------------------------------------------------------ lattices
struct D2Q9 {
  static constexpr int Q = 9, D = 2, ID = PYLBM_D2Q9;
  __host__ __device__ static constexpr int c(int i, int a) {
    constexpr int t[9][3] = {{0, 0, 0},  {1, 0, 0},  {0, 1, 0},   {-1, 0, 0}, {0, -1, 0},
                             {1, 1, 0},  {-1, 1, 0}, {-1, -1, 0}, {1, -1, 0}};
    return t[i][a];
  }
  __host__ __device__ static constexpr int opp(int i) {
    constexpr int t[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
    return t[i];
  }
  __host__ __device__ static constexpr double w(int i) {
    return i == 0 ? 4.0 / 9.0 : (i < 5 ? 1.0 / 9.0 : 1.0 / 36.0);
  }
};

struct D3Q19 {
  static constexpr int Q = 19, D = 3, ID = PYLBM_D3Q19;
  __host__ __device__ static constexpr int c(int i, int a) {
    constexpr int t[19][3] = {{0, 0, 0},   {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},
                              {0, 0, 1},   {0, 0, -1}, {1, 1, 0},  {-1, -1, 0}, {1, -1, 0},
                              {-1, 1, 0},  {1, 0, 1},  {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1},
                              {0, 1, 1},   {0, -1, -1}, {0, 1, -1}, {0, -1, 1}};
    return t[i][a];
  }
  __host__ __device__ static constexpr int opp(int i) {
    return i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1);
  }
  __host__ __device__ static constexpr double w(int i) {
    return i == 0 ? 1.0 / 3.0 : (i < 7 ? 1.0 / 18.0 : 1.0 / 36.0);
  }
};

// D2Q9 Lallemand-Luo basis, rows (rho, e, eps, jx, qx, jy, qy, pxx, pxy).
__host__ __device__ constexpr int mrt_m(int k, int i) {
  constexpr int t[9][9] = {{1, 1, 1, 1, 1, 1, 1, 1, 1},      {-4, -1, -1, -1, -1, 2, 2, 2, 2},
                           {4, -2, -2, -2, -2, 1, 1, 1, 1},  {0, 1, 0, -1, 0, 1, -1, -1, 1},
                           {0, -2, 0, 2, 0, 1, -1, -1, 1},   {0, 0, 1, 0, -1, 1, 1, -1, -1},
                           {0, 0, -2, 0, 2, 1, 1, -1, -1},   {0, 1, -1, 1, -1, 0, 0, 0, 0},
                           {0, 0, 0, 0, 0, 1, -1, 1, -1}};
  return t[k][i];
}
__host__ __device__ constexpr int mrt_norm(int k) {
  constexpr int t[9] = {9, 36, 36, 6, 12, 6, 12, 4, 4};
  return t[k];
}

// ------------------------------------------------------------------------ storage codecs
template <int ST> struct Codec;

template <> struct Codec<PYLBM_F32> {
  using T = float;
  static constexpr bool shifted = false;
  __device__ static __forceinline__ float ld(T v) { return v; }
  __device__ static __forceinline__ T st(float v) { return v; }
};

// fp32 holding the deviation f - w_i: same bytes as PYLBM_F32, but the small part of f that
// the collision relaxes keeps ~7 significant digits (used by the differentiable rollouts).
template <> struct Codec<PYLBM_F32S> {
  using T = float;
  static constexpr bool shifted = true;
  __device__ static __forceinline__ float ld(T v) { return v; }
  __device__ static __forceinline__ T st(float v) { return v; }
};

template <> struct Codec<PYLBM_F16> {
  using T = _Float16;
  static constexpr bool shifted = true;
  __device__ static __forceinline__ float ld(T v) { return (float)v; }
  // v_cvt_f16_f32 rounds to nearest even in the default mode.
  __device__ static __forceinline__ T st(float v) { return (_Float16)v; }
};

template <> struct Codec<PYLBM_BF16> {
  using T = uint16_t;
  static constexpr bool shifted = true;
  __device__ static __forceinline__ float ld(T v) {
    return __uint_as_float(((uint32_t)v) << 16);
  }
  // float -> bfloat16, round to nearest even (gfx950 converts in hardware).
  __device__ static __forceinline__ T st(float v) {
    __bf16 b = (__bf16)v;
    return __builtin_bit_cast(uint16_t, b);
  }
};

// ------------------------------------------------------------------------ kernel args
// Division by a run-time constant d on the scalar unit: floor(n / d) = (n * m) >> k with
// k = 32 + ceil(log2 d), m = ceil(2^k / d), exact for every n < 2^31 (the error term
// n (m - 2^k/d) / 2^k stays below 1/(2d)).  GPUs have no integer divide; without this,
// the row / column split of the block index would be computed on the vector unit.
struct FastDiv {
  uint64_t m;
  uint32_t k;
  __host__ __device__ uint32_t div(uint32_t n) const { return (uint32_t)(((uint64_t)n * m) >> k); }
};

inline FastDiv make_fastdiv(uint32_t d) {
  uint32_t l = 0;
  while (((uint64_t)1 << l) < d) ++l;
  const uint32_t k = 32 + l;
  const unsigned __int128 p = (unsigned __int128)1 << k;
  return FastDiv{(uint64_t)((p + d - 1) / d), k};
}

struct KArgs {
  int nx, ny, nz;
  int N;       // nodes per plane (< 2^31)
  int nseg;    // workgroups per row along x
  int nblk;    // total workgroups
  int rows;    // ny * nz
  int sw_log2;   // log2 of the lanes one row segment takes (8, or 7 / 6 for short rows)
  int rpw_log2;  // log2 of the rows one workgroup serves (0, or 1 / 2 for short rows)
  int per;     // bit a: axis a periodic
  int order;   // workgroup order over the XCDs (host policy, xcd_chunk)
  int chunk;   // consecutive workgroups an XCD takes per turn (0: plain round-robin)
  FastDiv div_nseg, div_ny, div_chunk;
  float ulid;
  float fx, fy, fz;
  float r0, r1, r2, r3;
  const float* ff;  // per-cell force (d, N) or nullptr
  // general geometry (k_step_geo only)
  const void* links;   // per-node solid-source mask or nullptr
  const float* inlet;  // u_x(y) of the velocity inlet on the x = 0 wall, or nullptr
  int outlet;          // 1: pressure outlet on the last column
};

// Workgroups are dispatched round-robin over the 8 XCDs, so workgroup b runs on XCD b % 8.
// With chunk C, XCD x takes C consecutive work items per turn: dispatch slot b maps to
// item ((i / C) * 8 + x) * C + i % C, with x = b % 8 and i = b / 8.  C = nblk / 8 gives
// every XCD its own slab (neighbouring segments and rows share one L2); small C keeps the
// eight XCDs sweeping one window of rows together; C = 0 leaves the dispatch order.
__device__ __forceinline__ int remap_block(int b, const KArgs& a) {
  if (a.chunk == 0) return b;
  const int x = b & 7, i = b >> 3;
  const int turn = (int)a.div_chunk.div((uint32_t)i);
  return (turn * 8 + x) * a.chunk + (i - turn * a.chunk);
}

// Host policy for the order (measured on MI355X, same-box A/B): states that fit in the
// 256 MiB Infinity Cache -> slabs (the 1024 x 256 adjoint channel ran 30% slower without);
// large D3Q19 states -> chunks of 16 (512^3 TRT: fp32 +2%, fp16 +6%, AA fp16 +4% over
// slabs, bf16 -1%); large D2Q9 states -> runs of 2 workgroups (16 at two nodes per lane):
// 4096^2 BGK over plain round-robin fp16 +1.9%, bf16 +2.1%, AA fp32 +2.2%, fp32 +0.3%,
// geometry fp32 -0.9%; round-robin itself beat slabs by 2-9%, and runs of 4 or 16
// workgroups at one node per lane lose 1-10%.  The read traffic seen by the L2 grows by
// the cache lines at workgroup edges that two XCDs both fetch: at 4096^2 fp32 reads are
// 1.083x the model under round-robin, 1.042x under runs of 2, 1.0002x under slabs.
// kOrderRun: runs of 2 workgroups per XCD turn at one node per lane, 16 at more.
enum XcdOrder { kOrderNone = 0, kOrderSlabs = 1, kOrderChunk16 = 2, kOrderRun = 3 };

inline void set_xcd_order(KArgs& a, int order, int npl = 1) {
  a.order = order;
  a.chunk = 0;
  const int run = order == kOrderChunk16 ? 16 : (order == kOrderRun ? (npl == 1 ? 2 : 16) : 0);
  if (run > 0 && a.nblk % (8 * run) == 0) a.chunk = run;
  else if (order != kOrderNone && a.nblk % 8 == 0) a.chunk = a.nblk / 8;
  a.div_chunk = make_fastdiv((uint32_t)(a.chunk > 0 ? a.chunk : 1));
}

// ------------------------------------------------------------------------ collision
// Per-node collision in fp32, written on deviations from the rest weights:
// h_i = f_i - w_i in, p_i - w_i out.  With rho = 1 + sum h, u = (sum c h + F/2) / rho and
// feq_i - w_i = w_i (drho + rho (3 cu + 4.5 cu^2 - 1.5 u^2)), every operator below is the
// reference's (pylbm/core.py:203-247) with f and feq replaced by their deviations (the
// weights cancel in f - feq and sum_i c_i w_i = 0), so the small non-equilibrium part is
// never formed as a difference of two O(w) numbers.  F: Guo force at the node; FORCE =
// false drops every Guo term at compile time (the no-force kernels).  1/rho uses the
// hardware reciprocal (v_rcp_f32, 1 ulp).
template <class L, int COLL, bool FORCE>
__device__ __forceinline__ void collide_node(const float (&h)[L::Q], float (&d)[L::Q],
                                             const float F[3], float r0, float r1, float r2,
                                             float r3) {
  constexpr int Q = L::Q;
  float drho = 0.f, j0 = 0.f, j1 = 0.f, j2 = 0.f;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    drho += h[i];
    if (L::c(i, 0)) j0 += L::c(i, 0) * h[i];
    if (L::c(i, 1)) j1 += L::c(i, 1) * h[i];
    if (L::D == 3 && L::c(i, 2)) j2 += L::c(i, 2) * h[i];
  }
  const float rho = 1.0f + drho;
  const float inv = __builtin_amdgcn_rcpf(rho);
  const float ux = (FORCE ? j0 + 0.5f * F[0] : j0) * inv;
  const float uy = (FORCE ? j1 + 0.5f * F[1] : j1) * inv;
  const float uz = (L::D == 3) ? (FORCE ? j2 + 0.5f * F[2] : j2) * inv : 0.0f;
  const float usq = ux * ux + uy * uy + uz * uz;
  const float uf = FORCE ? ux * F[0] + uy * F[1] + uz * F[2] : 0.0f;
  float neq[Q], S[Q];  // neq = f - feq = h - (feq - w)
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const float cu = L::c(i, 0) * ux + L::c(i, 1) * uy + L::c(i, 2) * uz;
    const float w = (float)L::w(i);
    const float poly = cu * (3.0f + 4.5f * cu) - 1.5f * usq;
    neq[i] = h[i] - w * (drho + rho * poly);
    const float cf = L::c(i, 0) * F[0] + L::c(i, 1) * F[1] + L::c(i, 2) * F[2];
    S[i] = FORCE ? w * (3.0f * (cf - uf) + 9.0f * cu * cf) : 0.0f;
  }
  if constexpr (COLL == PYLBM_BGK) {
    const float g = 1.0f - 0.5f * r0;
#pragma unroll
    for (int i = 0; i < Q; ++i) d[i] = FORCE ? h[i] - r0 * neq[i] + g * S[i] : h[i] - r0 * neq[i];
  } else if constexpr (COLL == PYLBM_TRT) {
    const float gp = 1.0f - 0.5f * r0, gm = 1.0f - 0.5f * r1;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int o = L::opp(i);
      const float np_ = 0.5f * (neq[i] + neq[o]);
      const float nm_ = 0.5f * (neq[i] - neq[o]);
      const float sp_ = 0.5f * (S[i] + S[o]), sm_ = 0.5f * (S[i] - S[o]);
      d[i] = FORCE ? h[i] - r0 * np_ - r1 * nm_ + gp * sp_ + gm * sm_ : h[i] - r0 * np_ - r1 * nm_;
    }
  } else {
    static_assert(COLL != PYLBM_MRT || L::Q == 9, "MRT is D2Q9 only");
    const float s[9] = {0.f, r1, r2, 0.f, r3, 0.f, r3, r0, r0};
    float dm[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      // the conserved moments (rho, jx, jy: rate 0) change only through the force; their
      // relaxation term is an exact zero, so it is not formed (nor, without a force, the
      // moment at all: 45 of the ~270 vector instructions of a node)
      const bool conserved = k == 0 || k == 3 || k == 5;
      float mne = 0.f, ms = 0.f;
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        if (mrt_m(k, i) == 0) continue;
        if (!conserved) mne += (float)mrt_m(k, i) * neq[i];
        if (FORCE) ms += (float)mrt_m(k, i) * S[i];
      }
      const float rn = (float)(1.0 / mrt_norm(k));
      if (conserved) dm[k] = FORCE ? ms * rn : 0.0f;
      else dm[k] = FORCE ? (-s[k] * mne + (1.0f - 0.5f * s[k]) * ms) * rn : -s[k] * mne * rn;
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      float acc = h[i];
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        if (mrt_m(k, i) == 0 || (!FORCE && (k == 0 || k == 3 || k == 5))) continue;
        acc += (float)mrt_m(k, i) * dm[k];
      }
      d[i] = acc;
    }
  }
}

}  // namespace plbm
