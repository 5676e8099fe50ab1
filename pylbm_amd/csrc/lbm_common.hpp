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
  int reverse; // 1: work items in reverse order (the sweep direction of this launch)
  int resident; // 1: the state fits in the 256 MiB Infinity Cache (slab order, temporal
                // stores, alternating sweeps)
  FastDiv div_nseg, div_ny, div_chunk;
  float ulid;
  float fx, fy, fz;
  float r0, r1, r2, r3;
  const float* ff;  // per-cell force (d, N) or nullptr
  // general geometry (k_step_geo only)
  const void* links;   // per-node solid-source mask or nullptr
  const float* inlet;  // u_x(y) of the velocity inlet on the x = 0 wall, or nullptr
  int outlet;          // 1: pressure outlet on the last column
  // moment records (k_step with REC, k_adjoint with REC): N x 4 fp32, node-major
  float* rec;
  // group flags of the link mask (k_step_geo with GF): one word per 64-node group of a row
  const uint32_t* lgroups;
  int gpr;  // groups per row, ceil(nx / 64)
};

// Workgroups are dispatched round-robin over the 8 XCDs, so workgroup b runs on XCD b % 8.
// With chunk C, XCD x takes C consecutive work items per turn: dispatch slot b maps to
// item ((i / C) * 8 + x) * C + i % C, with x = b % 8 and i = b / 8.  C = nblk / 8 gives
// every XCD its own slab (neighbouring segments and rows share one L2); small C keeps the
// eight XCDs sweeping one window of rows together; C = 0 leaves the dispatch order.
//
// reverse = 1 runs each XCD's sequence of work items backwards (when nblk is a multiple of
// 8, as it always is with chunk > 0, every XCD keeps its own items, so a slab stays with
// the L2 that holds it; otherwise the whole order is reversed).  Multi-step loops
// alternate the direction where it pays (the AA pattern, states resident in the 256 MiB
// Infinity Cache): a launch then starts on the rows the previous one wrote last, which
// the caches still hold.
__device__ __forceinline__ int remap_block(int b, const KArgs& a) {
  if (a.reverse) {
    if ((a.nblk & 7) == 0) b = ((((a.nblk >> 3) - 1) - (b >> 3)) << 3) | (b & 7);
    else b = a.nblk - 1 - b;
  }
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
// k-th opposite pair (i, opp(i)) with i < opp(i), in increasing i.
template <class L>
__host__ __device__ constexpr int pair_lo(int k) {
  int n = 0;
  for (int i = 1; i < L::Q; ++i) {
    if (L::opp(i) > i) {
      if (n == k) return i;
      ++n;
    }
  }
  return 0;
}
// c_i . v with the zero components of c_i left out and the unit ones as plain adds and
// subtractions (the components of every lattice here are -1, 0 or 1).  The sum of the kept
// terms is the value of the full product c_i,x v_x + c_i,y v_y + c_i,z v_z, which adds only
// signed zeros to it, up to the sign of a zero result, and saves the vector instructions
// that multiply by zero (they cannot be folded away: 0 * x is not 0 for every float x).
template <class L>
__device__ __forceinline__ float cdot(int i, float x, float y, float z) {
  float s = 0.0f;
  bool first = true;
  auto term = [&](int c, float v) {
    if (c == 0) return;
    if (first) {
      s = c > 0 ? v : -v;
      first = false;
    } else {
      s = c > 0 ? s + v : s - v;
    }
  };
  term(L::c(i, 0), x);
  term(L::c(i, 1), y);
  if (L::D == 3) term(L::c(i, 2), z);
  return s;
}

// Per-node collision in fp32, written on deviations from the rest weights:
// h_i = f_i - w_i in, p_i - w_i out.  With rho = 1 + sum h, u = (sum c h + F/2) / rho and
// feq_i - w_i = w_i (drho + rho (3 cu + 4.5 cu^2 - 1.5 u^2)), every operator below is the
// reference's (pylbm/core.py:203-247) with f and feq replaced by their deviations (the
// weights cancel in f - feq and sum_i c_i w_i = 0), so the small non-equilibrium part is
// never formed as a difference of two O(w) numbers.  F: Guo force at the node; FORCE =
// false drops every Guo term at compile time (the no-force kernels).  1/rho uses the
// hardware reciprocal (v_rcp_f32, 1 ulp).
//
// BGK works per population.  TRT and MRT work on opposite pairs (collide_pairs), which
// needs about 30% fewer vector instructions than per-population TRT and MRT; for BGK,
// which is bandwidth bound, the per-population form is kept because it rounds less (the
// 128^2 cavity after 1000 steps: 2.3e-7 against 3.8e-7 in the pair form).
template <class L, bool FORCE>
__device__ __forceinline__ void collide_bgk(const float (&h)[L::Q], float (&d)[L::Q],
                                            const float F[3], float r0, float (&mo)[4]) {
  // multiply-adds fuse within a statement only (as in collide_pairs), so the rounding does
  // not depend on the kernel this is inlined into
#pragma clang fp contract(on)
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
  mo[0] = rho;
  mo[1] = ux;
  mo[2] = uy;
  mo[3] = uz;
  const float usq = ux * ux + uy * uy + uz * uz;
  const float uf = FORCE ? ux * F[0] + uy * F[1] + uz * F[2] : 0.0f;
  const float g = 1.0f - 0.5f * r0;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const float cu = cdot<L>(i, ux, uy, uz);
    const float w = (float)L::w(i);
    const float poly = cu * (3.0f + 4.5f * cu) - 1.5f * usq;
    const float neq = h[i] - w * (drho + rho * poly);  // f - feq = h - (feq - w)
    const float cf = cdot<L>(i, F[0], F[1], F[2]);
    const float S = FORCE ? w * (3.0f * (cf - uf) + 9.0f * cu * cf) : 0.0f;
    d[i] = FORCE ? h[i] - r0 * neq + g * S : h[i] - r0 * neq;
  }
}

template <class L, int COLL, bool FORCE>
__device__ __forceinline__ void collide_pairs(const float (&h)[L::Q], float (&d)[L::Q],
                                              const float F[3], float r0, float r1, float r2,
                                              float r3, float (&mo)[4]) {
  // multiply-adds fuse within a statement only, never across statements, so the rounding
  // does not depend on the kernel this is inlined into (box or geometry kernel, one or
  // several nodes per lane, AB or AA pattern)
#pragma clang fp contract(on)
  constexpr int Q = L::Q;
  constexpr int NP = (Q - 1) / 2;
  // pair sums and differences of the deviations; the moments come from them:
  // drho = h_0 + sum (h_i + h_o), j = sum c_i (h_i - h_o)
  float hs[NP], hd[NP];
  float drho = h[0], j0 = 0.f, j1 = 0.f, j2 = 0.f;
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    const int i = pair_lo<L>(k), o = L::opp(i);
    hs[k] = h[i] + h[o];
    hd[k] = h[i] - h[o];
    drho += hs[k];
    if (L::c(i, 0)) j0 += L::c(i, 0) * hd[k];
    if (L::c(i, 1)) j1 += L::c(i, 1) * hd[k];
    if (L::D == 3 && L::c(i, 2)) j2 += L::c(i, 2) * hd[k];
  }
  const float rho = 1.0f + drho;
  const float inv = __builtin_amdgcn_rcpf(rho);
  const float ux = (FORCE ? j0 + 0.5f * F[0] : j0) * inv;
  const float uy = (FORCE ? j1 + 0.5f * F[1] : j1) * inv;
  const float uz = (L::D == 3) ? (FORCE ? j2 + 0.5f * F[2] : j2) * inv : 0.0f;
  mo[0] = rho;
  mo[1] = ux;
  mo[2] = uy;
  mo[3] = uz;
  const float usq = ux * ux + uy * uy + uz * uz;
  const float uf = FORCE ? ux * F[0] + uy * F[1] + uz * F[2] : 0.0f;
  {
    // Opposite pairs (i, o = opp(i)).  With c_o = -c_i (cu_o = -cu_i,
    // cf_o = -cf_i) the pair sums and differences of f - feq and of the Guo term have
    // closed forms, so neither is formed per population:
    //   sn = neq_i + neq_o = h_i + h_o - w (2 drho + rho (9 cu^2 - 3 u^2))
    //   dn = neq_i - neq_o = h_i - h_o - 6 w rho cu
    //   ss = S_i + S_o     = w (18 cu cf - 6 uf)
    //   ds = S_i - S_o     = 6 w cf
    // The rest population (c = 0) has neq_0 = h_0 - w_0 (drho - 1.5 rho u^2) and
    // S_0 = -3 w_0 uf.
    float sn[NP], ds[NP], dn[NP], ss[NP];
    const float drho2 = 2.0f * drho, usq3 = 3.0f * usq;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const int i = pair_lo<L>(k);
      const float cu = cdot<L>(i, ux, uy, uz);
      const float w = (float)L::w(i);
      sn[k] = hs[k] - w * (drho2 + rho * (9.0f * cu * cu - usq3));
      dn[k] = hd[k] - (6.0f * w) * rho * cu;
      if (FORCE) {
        const float cf = cdot<L>(i, F[0], F[1], F[2]);
        ss[k] = w * (18.0f * cu * cf - 6.0f * uf);
        ds[k] = (6.0f * w) * cf;
      } else {
        ss[k] = ds[k] = 0.0f;
      }
    }
    const float w0 = (float)L::w(0);
    const float n0 = h[0] - w0 * (drho - 1.5f * rho * usq);
    const float S0 = FORCE ? -3.0f * w0 * uf : 0.0f;
    if constexpr (COLL == PYLBM_TRT) {
      // d_{i,o} = h_{i,o} - a -/+ b, a = (s+ sn - g+ ss) / 2, b = (s- dn - g- ds) / 2
      const float rm = r1;
      const float gp = 1.0f - 0.5f * r0, gm = 1.0f - 0.5f * rm;
      const float hp = 0.5f * r0, hm = 0.5f * rm, hgp = 0.5f * gp, hgm = 0.5f * gm;
      d[0] = FORCE ? h[0] - r0 * n0 + gp * S0 : h[0] - r0 * n0;
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        const int i = pair_lo<L>(k), o = L::opp(i);
        if (FORCE) {
          d[i] = h[i] - hp * sn[k] + hgp * ss[k] - hm * dn[k] + hgm * ds[k];
          d[o] = h[o] - hp * sn[k] + hgp * ss[k] + hm * dn[k] - hgm * ds[k];
        } else {
          d[i] = h[i] - hp * sn[k] - hm * dn[k];
          d[o] = h[o] - hp * sn[k] + hm * dn[k];
        }
      }
    } else {
      // D2Q9 Lallemand-Luo moments of neq (and of S) from the pair parts; the pairs are
      // (1, 3), (2, 4), (5, 7), (6, 8):
      //   e   = -4 n0 - (S13 + S24) + 2 (S57 + S68)   eps = 4 n0 - 2 (S13 + S24) + (S57 + S68)
      //   qx  = -2 D13 + D57 - D68                      qy  = -2 D24 + D57 + D68
      //   pxx = S13 - S24                               pxy = S57 - S68
      // (rows of mrt_m; rho, jx, jy are conserved: rate 0, they change only through the
      // force: rho(S) = S0 + sum ss, jx(S) = ds13 + ds57 - ds68, jy(S) = ds24 + ds57 + ds68).
      // dm_k = (-s_k m_k(neq) + (1 - s_k / 2) m_k(S)) / |M_k|^2, then p = h + M^T dm:
      //   p0 = h0 - 4 dm_e + 4 dm_eps (+ dm_rho)
      //   p1,3 = h - (dm_e + 2 dm_eps) + dm_pxx (+ dm_rho) +/- (-2 dm_qx (+ dm_jx))
      //   p2,4 = h - (dm_e + 2 dm_eps) - dm_pxx (+ dm_rho) +/- (-2 dm_qy (+ dm_jy))
      //   p5,7 = h + (2 dm_e + dm_eps) + dm_pxy (+ dm_rho) +/- (dm_qx + dm_qy (+ dm_jx + dm_jy))
      //   p6,8 = h + (2 dm_e + dm_eps) - dm_pxy (+ dm_rho) +/- (-dm_qx + dm_qy (- dm_jx + dm_jy))
      static_assert(COLL != PYLBM_MRT || L::Q == 9, "MRT is D2Q9 only");
      static_assert(L::Q != 9 || (pair_lo<L>(0) == 1 && pair_lo<L>(1) == 2 &&
                                  pair_lo<L>(2) == 5 && pair_lo<L>(3) == 6 && L::opp(1) == 3 &&
                                  L::opp(2) == 4 && L::opp(5) == 7 && L::opp(6) == 8),
                    "D2Q9 pair order");
      // pair slots k: 0 = (1,3), 1 = (2,4), 2 = (5,7), 3 = (6,8)
      const float sA = sn[0] + sn[1], sB = sn[2] + sn[3];
      const float m_e = -4.0f * n0 - sA + 2.0f * sB;
      const float m_eps = 4.0f * n0 - 2.0f * sA + sB;
      const float m_qx = -2.0f * dn[0] + dn[2] - dn[3];
      const float m_qy = -2.0f * dn[1] + dn[2] + dn[3];
      const float m_pxx = sn[0] - sn[1];
      const float m_pxy = sn[2] - sn[3];
      // rates: s_e = r1, s_eps = r2, s_q = r3, s_nu = r0 (pxx, pxy); -s_k / |M_k|^2 and
      // (1 - s_k / 2) / |M_k|^2 are uniform factors
      const float ke = -r1 * (1.0f / 36.0f), keps = -r2 * (1.0f / 36.0f);
      const float kq = -r3 * (1.0f / 12.0f), kp = -r0 * 0.25f;
      float dm_e, dm_eps, dm_qx, dm_qy, dm_pxx, dm_pxy;
      float dm_rho = 0.0f, dm_jx = 0.0f, dm_jy = 0.0f;
      if (FORCE) {
        const float ge = (1.0f - 0.5f * r1) * (1.0f / 36.0f);
        const float geps = (1.0f - 0.5f * r2) * (1.0f / 36.0f);
        const float gq = (1.0f - 0.5f * r3) * (1.0f / 12.0f);
        const float gpx = (1.0f - 0.5f * r0) * 0.25f;
        const float tA = ss[0] + ss[1], tB = ss[2] + ss[3];
        dm_e = ke * m_e + ge * (-4.0f * S0 - tA + 2.0f * tB);
        dm_eps = keps * m_eps + geps * (4.0f * S0 - 2.0f * tA + tB);
        dm_qx = kq * m_qx + gq * (-2.0f * ds[0] + ds[2] - ds[3]);
        dm_qy = kq * m_qy + gq * (-2.0f * ds[1] + ds[2] + ds[3]);
        dm_pxx = kp * m_pxx + gpx * (ss[0] - ss[1]);
        dm_pxy = kp * m_pxy + gpx * (ss[2] - ss[3]);
        dm_rho = (S0 + tA + tB) * (1.0f / 9.0f);
        dm_jx = (ds[0] + ds[2] - ds[3]) * (1.0f / 6.0f);
        dm_jy = (ds[1] + ds[2] + ds[3]) * (1.0f / 6.0f);
      } else {
        dm_e = ke * m_e;
        dm_eps = keps * m_eps;
        dm_qx = kq * m_qx;
        dm_qy = kq * m_qy;
        dm_pxx = kp * m_pxx;
        dm_pxy = kp * m_pxy;
      }
      const float E1 = dm_e + 2.0f * dm_eps, E2 = 2.0f * dm_e + dm_eps;
      const float P13 = dm_pxx - E1 + dm_rho, Q13 = dm_jx - 2.0f * dm_qx;
      const float P24 = -dm_pxx - E1 + dm_rho, Q24 = dm_jy - 2.0f * dm_qy;
      const float P57 = E2 + dm_pxy + dm_rho, Q57 = (dm_qx + dm_qy) + (dm_jx + dm_jy);
      const float P68 = E2 - dm_pxy + dm_rho, Q68 = (dm_qy - dm_qx) + (dm_jy - dm_jx);
      d[0] = h[0] + 4.0f * (dm_eps - dm_e) + dm_rho;
      d[1] = h[1] + P13 + Q13;
      d[3] = h[3] + P13 - Q13;
      d[2] = h[2] + P24 + Q24;
      d[4] = h[4] + P24 - Q24;
      d[5] = h[5] + P57 + Q57;
      d[7] = h[7] + P57 - Q57;
      d[6] = h[6] + P68 + Q68;
      d[8] = h[8] + P68 - Q68;
    }
  }
}

// mo receives the node's density and velocity as the collision used them: (rho, u_x, u_y,
// u_z) with u = (sum c h + F/2) / rho (u_z = 0 in 2D); kernels that do not record them let
// the compiler drop the copies.
template <class L, int COLL, bool FORCE>
__device__ __forceinline__ void collide_node(const float (&h)[L::Q], float (&d)[L::Q],
                                             const float F[3], float r0, float r1, float r2,
                                             float r3, float (&mo)[4]) {
  if constexpr (COLL == PYLBM_BGK)
    collide_bgk<L, FORCE>(h, d, F, r0, mo);
  else
    collide_pairs<L, COLL, FORCE>(h, d, F, r0, r1, r2, r3, mo);
}

template <class L, int COLL, bool FORCE>
__device__ __forceinline__ void collide_node(const float (&h)[L::Q], float (&d)[L::Q],
                                             const float F[3], float r0, float r1, float r2,
                                             float r3) {
  float mo[4];
  collide_node<L, COLL, FORCE>(h, d, F, r0, r1, r2, r3, mo);
}

}  // namespace plbm
