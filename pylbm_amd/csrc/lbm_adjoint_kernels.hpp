// Discrete adjoint of the fused two-buffer step (BGK / TRT with Guo forcing), for gradients
// of a loss of the final velocity field with respect to a per-cell Guo force F(x), tau and
// the initial velocity.  The reference has no gradient path; the float64 cross-checks are a
// torch-autograd restatement of the same step and central finite differences.
//
// Forward step at node x (pull, as in the step kernels):
//   g_i(x) = f_i(x - c_i), or f_opp(i)(x) + lid_i on a link crossing a wall
//   p(x)   = g - s+ P+ (g - feq(g)) - s- P- (g - feq(g)) + (1 - s+/2) P+ S + (1 - s-/2) P- S
// with P+- the (anti)symmetric projectors over opposite pairs (BGK: s+ = s- = omega), rho,
// u = (sum c g + F/2)/rho, feq_i = w_i rho (1 + 3 cu + 4.5 cu^2 - 1.5 u^2) and
// S_i = w_i (3 (c_i - u).F + 9 (c_i.u)(c_i.F)).
//
// Adjoint of one step, per node x, for lam = dL/dp(x):
//   lam'  = (s+ P+ + s- P-) lam,   lam'' = (g+ P+ + g- P-) lam,   g+- = 1 - s+-/2
//   A     = sum_i lam'_i feq_i / rho
//   B_a   = sum_i lam'_i dfeq_i/du_a + lam''_i dS_i/du_a
//   mu_k  = lam_k - lam'_k + A + sum_a B_a (c_ka - u_a) / rho          (= J^T lam)
// and mu_k is pushed back along the link it was pulled over: to (k, x - c_k), or to
// (opp(k), x) on a wall link.  Pull streaming with bounce-back is a permutation, so every
// slot of the previous adjoint state is written exactly once (no atomics).  The lid term is
// a constant and drops out.  Per-cell force and rate sensitivities:
//   dL/dF_b(x) += B_b / (2 rho) + sum_i lam''_i w_i (3 (c_ib - u_b) + 9 cu_i c_ib)
//   dL/ds+-    += -sum_i lam+-_i (g_i - feq_i + S_i / 2)      (lam+- = P+- lam)
#pragma once

#include "lbm_access.hpp"
#include "lbm_adjoint.hpp"


namespace plbm {
namespace adj {


// Per-node forward quantities: the moments of the pulled deviations h_i = g_i - w_i (exact
// in fp32 for g near w), summed in the order of the forward collision.
template <class L, bool FORCE>
__device__ __forceinline__ void pulled_moments(const float (&h)[L::Q], const float F[3],
                                               float& drho, float& rho, float& inv,
                                               float (&u)[3]) {
#pragma clang fp contract(on)
  constexpr int Q = L::Q, D = L::D;
  float j[3] = {0.f, 0.f, 0.f};
  drho = 0.f;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    drho += h[i];
#pragma unroll
    for (int a = 0; a < D; ++a)
      if (L::c(i, a)) j[a] += L::c(i, a) * h[i];
  }
  rho = 1.0f + drho;
  inv = __builtin_amdgcn_rcpf(rho);
  u[0] = u[1] = u[2] = 0.f;
#pragma unroll
  for (int a = 0; a < D; ++a) u[a] = (FORCE ? j[a] + 0.5f * F[a] : j[a]) * inv;
}

// The transposed collision Jacobian mu = J_C^T lam at one node, from rho, u (and F), in
// moment form.  lam+- = P+- lam, lam' = s+ lam+ + s- lam-, lam'' = g+ lam+ + g- lam-
// (BGK: s+ = s- = omega).  With the weighted moments of lam over opposite pairs
// (s_p = lam_i + lam_o, d_p = lam_i - lam_o):
//   m0 = sum w lam,  m1_a = sum w c_a lam,  M2_ab = sum w c_a c_b lam
// (m0 and M2 see lam+ only, m1 lam- only), the sums of the adjoint become
//   A    = sum lam' w (1 + poly) = s+ (m0 (1 - 1.5 u^2) + 4.5 u.M2.u) + 3 s- u.m1
//   B_a  = 3 rho (s- m1_a + 3 s+ (M2 u)_a - u_a s+ m0)
//          [+ 9 g+ (M2 F)_a - 3 F_a g+ m0                          with a force]
//   mu_k = lam_k - lam'_k + (A - B.u / rho) + c_k.B / rho
//   dF_a = B_a / (2 rho) + 3 g- m1_a + 9 g+ (M2 u)_a - 3 u_a g+ m0   (force only)
// poly = 3 cu + 4.5 cu^2 - 1.5 u^2.  About a third of the per-population form's vector
// instructions: the 10 moments take 63 multiply-adds for D3Q19, the pushes two per pair.
template <class L, int COLL, bool FORCE>
__device__ __forceinline__ void adjoint_core(float rho, float inv, const float (&u)[3],
                                             const float F[3], const float (&lam)[L::Q],
                                             float sp, float sm, float (&mu)[L::Q],
                                             float dF[3]) {
  // multiply-adds fuse within a statement only, so every adjoint kernel (two-buffer,
  // in place, two steps per launch) rounds every node identically
#pragma clang fp contract(on)
  constexpr int Q = L::Q, D = L::D, NP = (Q - 1) / 2;
  const float s_p = sp, s_m = (COLL == PYLBM_BGK) ? sp : sm;
  float ps[NP], pd[NP];
  float m0 = (float)L::w(0) * lam[0];
  float m1[3] = {0.f, 0.f, 0.f};
  float M2[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    const int i = pair_lo<L>(k), o = L::opp(i);
    ps[k] = lam[i] + lam[o];
    pd[k] = lam[i] - lam[o];
    const float w = (float)L::w(i);
    m0 += w * ps[k];
#pragma unroll
    for (int a = 0; a < D; ++a) {
      if (L::c(i, a) == 0) continue;
      m1[a] += (L::c(i, a) > 0 ? w : -w) * pd[k];
#pragma unroll
      for (int b = a; b < D; ++b)
        if (L::c(i, b) != 0) M2[a][b] += (L::c(i, a) * L::c(i, b) > 0 ? w : -w) * ps[k];
    }
  }
  float M2u[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int b = 0; b < D; ++b) M2u[a] += (a <= b ? M2[a][b] : M2[b][a]) * u[b];
  float uM2u = 0.f, um1 = 0.f;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    uM2u += u[a] * M2u[a];
    um1 += u[a] * m1[a];
  }
  const float usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const float A = s_p * (m0 * (1.0f - 1.5f * usq) + 4.5f * uM2u) + (3.0f * s_m) * um1;
  const float P = s_p * m0;
  const float gp = 1.0f - 0.5f * s_p, gm = 1.0f - 0.5f * s_m;
  float Bn[3] = {0.f, 0.f, 0.f}, Bu = 0.f;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    float B = 3.0f * rho * ((s_m * m1[a] + (3.0f * s_p) * M2u[a]) - u[a] * P);
    if (FORCE) {
      float M2F = 0.f;
#pragma unroll
      for (int b = 0; b < D; ++b) M2F += (a <= b ? M2[a][b] : M2[b][a]) * F[b];
      B += 9.0f * (gp * M2F) - 3.0f * F[a] * (gp * m0);
    }
    Bn[a] = B * inv;
    Bu += Bn[a] * u[a];
  }
  const float K0 = A - Bu;
  const float hp = 0.5f * (1.0f - s_p), hm = 0.5f * (1.0f - s_m);
  mu[0] = (1.0f - s_p) * lam[0] + K0;
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    const int i = pair_lo<L>(k), o = L::opp(i);
    float cB = 0.f;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      if (L::c(i, a) > 0) cB += Bn[a];
      else if (L::c(i, a) < 0) cB -= Bn[a];
    }
    const float base = hp * ps[k] + K0;
    const float diff = hm * pd[k] + cB;
    mu[i] = base + diff;
    mu[o] = base - diff;
  }
#pragma unroll
  for (int a = 0; a < D; ++a)
    dF[a] = FORCE ? 0.5f * Bn[a] + ((3.0f * gm) * m1[a] + (9.0f * gp) * M2u[a]) -
                        3.0f * u[a] * (gp * m0)
                  : 0.0f;
}

// The rate sensitivities of one node: dL/ds+- += -sum_i lam+-_i r_i with
// r_i = g_i - feq_i + S_i / 2 (BGK: dL/domega into dsp).  r is the small non-equilibrium
// part; it is formed per population from deviations (h_i - (feq_i - w_i)) before it meets
// lam, so it keeps its digits (summing lam h and lam (feq - w) separately would cancel).
template <class L, int COLL, bool FORCE>
__device__ __forceinline__ void rate_sensitivity(const float (&h)[L::Q], float drho,
                                                 float rho, const float (&u)[3],
                                                 const float F[3], const float (&lam)[L::Q],
                                                 float& dsp, float& dsm) {
#pragma clang fp contract(on)
  constexpr int Q = L::Q;
  const float usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const float uf = FORCE ? u[0] * F[0] + u[1] * F[1] + u[2] * F[2] : 0.0f;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int o = L::opp(i);
    const float w = (float)L::w(i);
    const float cu = L::c(i, 0) * u[0] + L::c(i, 1) * u[1] + L::c(i, 2) * u[2];
    const float poly = cu * (3.0f + 4.5f * cu) - 1.5f * usq;
    float S = 0.0f;
    if (FORCE) {
      const float cf = L::c(i, 0) * F[0] + L::c(i, 1) * F[1] + L::c(i, 2) * F[2];
      S = w * (3.0f * (cf - uf) + 9.0f * cu * cf);
    }
    const float r = (h[i] - w * (drho + rho * poly)) + 0.5f * S;  // g - feq + S/2
    if (COLL == PYLBM_BGK) {
      dsp -= lam[i] * r;
    } else {
      dsp -= 0.5f * (lam[i] + lam[o]) * r;
      dsm -= 0.5f * (lam[i] - lam[o]) * r;
    }
  }
}

// The adjoint state element of an fp32 value: fp32 as it is, or rounded to fp16.  The empty
// asm keeps the value an fp32 register first, as the source writes it: otherwise the
// compiler may fold the last multiply-add and the conversion into one mixed-precision
// instruction that rounds once to fp16 (v_fma_mixlo_f16), in some kernels and not in
// others, and the in-place and two-buffer adjoints would differ in the last fp16 bit.
template <class LT>
__device__ __forceinline__ LT to_lam(float v) {
  if constexpr (sizeof(LT) == 2) asm volatile("" : "+v"(v));
  return (LT)v;
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// LT is the element type of the adjoint states lam_in / lam_out: float, or _Float16 (values
// rounded to nearest fp16 on store; every operation on them is fp32).
//
// PAT selects the layout of the adjoint states (the AA pattern run backwards):
//   PYLBM_AB       lam_in natural -> lam_out natural, two buffers (push to (i, x - c_i),
//                  or to the own slot opp(i) on a wall link);
//   PYLBM_AA_EVEN  in place: reads the own natural slots (i, x), writes mu_i to the own
//                  slot opp(i) (the push along c_i is left to the next step);
//   PYLBM_AA_ODD   in place: reads lam_i(x) from (opp(i), x + c_i), where the even step
//                  left it (from the own slot i when x + c_i is behind a wall: the
//                  bounced-back mu_opp(i)(x)), and pushes mu_i as PYLBM_AB does, which
//                  leaves the state natural again.
// In both AA steps a node writes exactly the slots it read, so the update in place has no
// race; a pair of them equals two two-buffer steps bit for bit.
//
// REC: the forward quantities come from the step's moment record (rho, u: 16 bytes per
// node, written by the recomputing forward step) instead of a pull of the stored state;
// f is then the record array.  The rate sensitivities need the pulled populations, so
// REC excludes RS.
template <class L, int COLL, int ST, int FM, int ADDR, class LT, int SP, int PAT, bool RS,
          bool REC = false>
__global__ void __launch_bounds__(kBlock)
k_adjoint(const KArgs a, const typename Codec<ST>::T* __restrict__ f, const LT* lam_in,
          LT* lam_out, float* __restrict__ dF, double* __restrict__ part) {
  static_assert(!(REC && RS), "the rate sensitivities need the pulled populations");
  using T = typename Codec<ST>::T;
  constexpr int Q = L::Q;
  constexpr uint32_t TS = sizeof(T);
  constexpr uint32_t LS = sizeof(LT);
  Lane l;
  const bool live = lane_setup<TS>(a, remap_block((int)blockIdx.x, a), l);
  float dsp = 0.f, dsm = 0.f;
  if (live) {
    const size_t N = (size_t)a.N;
    const char* fb = reinterpret_cast<const char*>(f);
    const char* lib = reinterpret_cast<const char*>(lam_in);
    char* lob = reinterpret_cast<char*>(lam_out);
    const size_t lplane = N * LS;      // adjoint-state plane
    const size_t fplane = N * 4u;      // fp32 force / dF plane
    const uint32_t q0 = l.n * LS, qm = l.nm * LS, qp = l.np * LS;  // adjoint-state lane offsets
    const uint32_t f0 = l.n * 4u;                                   // fp32 lane offset
    // every load (pulled f, own lambda, force, dF) is issued before the first wait
    float h[Q], lam[Q];
    T t[Q];
    float lid[Q];
    float mo[4];
    if constexpr (REC) {
      // read once, by this step only: non-temporal
      ld_rec_nt(reinterpret_cast<const float*>(f), l.n, mo);
    } else {
      // resident states (SP = 0): the forward state each step reads once streams past the
      // Infinity Cache (non-temporal loads), which keeps the adjoint state there
      gather_issue<L, ST, PYLBM_AB, ADDR, SP == 0 ? 2 : 0>(a, l, fb, N * TS, t, lid);
    }
    if constexpr (PAT == PYLBM_AA_ODD) {
      // lam_i(x) is mu_i(x + c_i), which the even step left in (opp(i), x + c_i); when
      // x + c_i lies behind a wall it is the bounced-back mu_opp(i)(x), left in the own
      // slot i (y/z walls are uniform and selected by value; x-wall lanes re-read their
      // own slots afterwards, in one divergent block, as the gather does)
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const int cx = L::c(i, 0);
        const RowLink r = row_link<L>(a, l.y, l.z, i, +1);
        lam[i] = (float)ld_pl<LT, ADDR>(lib, lplane, r.wall ? i : L::opp(i),
                                        r.wall ? 0 : r.delta * (ptrdiff_t)LS,
                                        r.wall ? q0 : (cx > 0 ? qp : (cx < 0 ? qm : q0)));
      }
      if (l.wall_lo || l.wall_hi) {
        LT own[Q];
#pragma unroll
        for (int i = 1; i < Q; ++i)
          if (L::c(i, 0) != 0) own[i] = ld_pl<LT, ADDR>(lib, lplane, i, 0, q0);
#pragma unroll
        for (int i = 1; i < Q; ++i)
          if ((L::c(i, 0) > 0 && l.wall_hi) || (L::c(i, 0) < 0 && l.wall_lo))
            lam[i] = (float)own[i];
      }
    } else {
#pragma unroll
      for (int i = 0; i < Q; ++i) lam[i] = (float)ld_pl<LT, ADDR>(lib, lplane, i, 0, q0);
    }
    constexpr bool FF = FM == 2;
    float F[3] = {a.fx, a.fy, a.fz};
    const char* ffb = reinterpret_cast<const char*>(a.ff);
    char* dFb = reinterpret_cast<char*>(dF);
    float dF0[3] = {0.f, 0.f, 0.f};
    if (FF) {
#pragma unroll
      for (int d = 0; d < L::D; ++d) F[d] = ld_pl<float, ADDR>(ffb, fplane, d, 0, f0);
#pragma unroll
      for (int d = 0; d < L::D; ++d) dF0[d] = ld_pl<float, ADDR>(dFb, fplane, d, 0, f0);
    }
    float mu[Q], dFn[3], drho, rho, inv, u[3];
    if constexpr (REC) {
      rho = mo[0];
      inv = __builtin_amdgcn_rcpf(rho);
      u[0] = mo[1];
      u[1] = mo[2];
      u[2] = mo[3];
    } else {
      gather_finish<L, ST, PYLBM_AB, ADDR>(l, fb, N * TS, t, lid, h);
      pulled_moments<L, FM != 0>(h, F, drho, rho, inv, u);
    }
    adjoint_core<L, COLL, FM != 0>(rho, inv, u, F, lam, a.r0, a.r1, mu, dFn);
    if (RS) rate_sensitivity<L, COLL, FM != 0>(h, drho, rho, u, F, lam, dsp, dsm);
    // push mu_i back along the link it was pulled over: to (i, x - c_i), or to the own
    // slot opp(i) on a wall link.  Waves with no lane on an x wall (every wave of a box
    // periodic in x) take a loop without per-lane branches; the per-lane wall choice costs
    // each direction a divergent branch and its exec-mask bookkeeping on the scalar unit.
    if constexpr (PAT == PYLBM_AA_EVEN) {
#pragma unroll
      for (int i = 0; i < Q; ++i) st_pl<SP, ADDR>(lob, lplane, L::opp(i), 0, q0, to_lam<LT>(mu[i]));
    } else if (__ballot(l.wall_lo || l.wall_hi) == 0) {
      // a y/z wall (uniform) selects the own slot opp(i) by value, without a branch
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const int cx = L::c(i, 0);
        const RowLink r = row_link<L>(a, l.y, l.z, i, -1);
        const bool own = i == 0 || r.wall;
        st_pl<SP, ADDR>(lob, lplane, own ? L::opp(i) : i, own ? 0 : r.delta * (ptrdiff_t)LS,
                        own ? q0 : (cx > 0 ? qm : (cx < 0 ? qp : q0)), to_lam<LT>(mu[i]));
      }
    } else {
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const LT m = to_lam<LT>(mu[i]);
        if (i == 0) {
          st_pl<SP, ADDR>(lob, lplane, 0, 0, q0, m);
          continue;
        }
        const int cx = L::c(i, 0);
        const RowLink r = row_link<L>(a, l.y, l.z, i, -1);
        if (r.wall) {
          st_pl<SP, ADDR>(lob, lplane, L::opp(i), 0, q0, m);
          continue;
        }
        const uint32_t off = cx > 0 ? qm : (cx < 0 ? qp : q0);
        if ((cx > 0 && l.wall_lo) || (cx < 0 && l.wall_hi))
          st_pl<SP, ADDR>(lob, lplane, L::opp(i), 0, q0, m);
        else
          st_pl<SP, ADDR>(lob, lplane, i, r.delta * (ptrdiff_t)LS, off, m);
      }
    }
    if (FF) {
#pragma unroll
      for (int d = 0; d < L::D; ++d) st_pl<0, ADDR>(dFb, fplane, d, 0, f0, dF0[d] + dFn[d]);
    }
  }
  if (!RS) return;
  // per-workgroup partial sums of dL/ds+ and dL/ds- (deterministic: one slot per group)
  __shared__ float red[2][kBlock / 64];
  dsp = wave_sum_f(dsp);
  dsm = wave_sum_f(dsm);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wv] = dsp;
    red[1][wv] = dsm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s0 = 0.0, s1 = 0.0;
    for (int k = 0; k < kBlock / 64; ++k) {
      s0 += red[0][k];
      s1 += red[1][k];
    }
    const size_t b = blockIdx.x;
    part[2 * b] += s0;
    part[2 * b + 1] += s1;
  }
}

// Terminal adjoint state of L(u(f)), u = (sum c f - F/2)/rho post-collision:
//   lam_k = sum_a G_a (c_ka - u_a) / rho,   dF_a = -G_a / (2 rho)   (G = dL/du)
template <class L, int ST, bool FF, class LT>
__global__ void __launch_bounds__(kBlock)
k_adjoint_seed(const KArgs a, const typename Codec<ST>::T* __restrict__ f,
               const float* __restrict__ G, LT* __restrict__ lam, float* __restrict__ dF) {
  using C = Codec<ST>;
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= a.N) return;
  const size_t N = (size_t)a.N;
  float fv[L::Q], rho = 0.f, j[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int i = 0; i < L::Q; ++i) {
    fv[i] = C::ld(f[(size_t)i * N + n]) + (C::shifted ? (float)L::w(i) : 0.0f);
    rho += fv[i];
#pragma unroll
    for (int d = 0; d < L::D; ++d) j[d] += L::c(i, d) * fv[i];
  }
  float F[3] = {a.fx, a.fy, a.fz};
  if (FF) {
#pragma unroll
    for (int d = 0; d < L::D; ++d) F[d] = a.ff[(size_t)d * N + n];
  }
  const float inv = 1.0f / rho;
  float u[3] = {0.f, 0.f, 0.f}, Gv[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int d = 0; d < L::D; ++d) {
    u[d] = (j[d] - 0.5f * F[d]) * inv;
    Gv[d] = G[(size_t)d * N + n];
  }
#pragma unroll
  for (int k = 0; k < L::Q; ++k) {
    float m = 0.f;
#pragma unroll
    for (int d = 0; d < L::D; ++d) m += Gv[d] * ((float)L::c(k, d) - u[d]);
    lam[(size_t)k * N + n] = (LT)(m * inv);
  }
  if (FF) {
#pragma unroll
    for (int d = 0; d < L::D; ++d) dF[(size_t)d * N + n] = -0.5f * Gv[d] * inv;
  }
}

// Gradient of the initial equilibrium f0 = feq(rho0, u0) (float64 fields) with respect to
// u0 and rho0, given lam0 = dL/df0.
template <class L, class LT>
__global__ void __launch_bounds__(kBlock)
k_adjoint_init(int N_, const double* __restrict__ rho0, const double* __restrict__ u0,
               const LT* __restrict__ lam, double* __restrict__ du, double* __restrict__ drho) {
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= N_) return;
  const size_t N = (size_t)N_;
  const double r = rho0 ? rho0[n] : 1.0;
  double u[3] = {0.0, 0.0, 0.0};
  if (u0) {
#pragma unroll
    for (int d = 0; d < L::D; ++d) u[d] = u0[(size_t)d * N + n];
  }
  const double usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  double gu[3] = {0.0, 0.0, 0.0}, gr = 0.0;
#pragma unroll
  for (int i = 0; i < L::Q; ++i) {
    const double l = (double)(float)lam[(size_t)i * N + n];
    const double cu = L::c(i, 0) * u[0] + L::c(i, 1) * u[1] + L::c(i, 2) * u[2];
    gr += l * L::w(i) * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq);
#pragma unroll
    for (int d = 0; d < L::D; ++d)
      gu[d] += l * L::w(i) * r * (3.0 * L::c(i, d) + 9.0 * cu * L::c(i, d) - 3.0 * u[d]);
  }
  if (du) {
#pragma unroll
    for (int d = 0; d < L::D; ++d) du[(size_t)d * N + n] = gu[d];
  }
  if (drho) drho[n] = gr;
}

template <class L, int COLL, int ST, class LT, int PAT, bool RS>
int adj_step_tpr(const KArgs& a, bool field, const void* f, const void* lin, void* lout,
                 float* dF, double* part, hipStream_t s) {
  using T = typename Codec<ST>::T;
  const dim3 grid((unsigned)a.nblk), block(kBlock);
  const bool uniform = a.fx != 0.0f || a.fy != 0.0f || a.fz != 0.0f;
  const T* fs = (const T*)f;
  const LT* li = (const LT*)lin;
  LT* lo = (LT*)lout;
  // one descriptor per array when a q-plane fp32 array (the largest the kernel can touch)
  // spans less than 4 GiB
  const bool small = (uint64_t)L::Q * (uint64_t)a.N * 4u < ((uint64_t)1 << 32);
  // SP: temporal adjoint-state stores when the forward state is resident in the Infinity
  // Cache (the next adjoint step re-reads them), non-temporal otherwise
#define PLBM_ADJ(FMv, ADDRv, SPv)                                                          \
  k_adjoint<L, COLL, ST, FMv, ADDRv, LT, SPv, PAT, RS><<<grid, block, 0, s>>>(a, fs, li, lo, \
                                                                            dF, part)
  if (small && a.resident) {
    if (field) PLBM_ADJ(2, 1, 0);
    else if (uniform) PLBM_ADJ(1, 1, 0);
    else PLBM_ADJ(0, 1, 0);
  } else if (small) {
    if (field) PLBM_ADJ(2, 1, 2);
    else if (uniform) PLBM_ADJ(1, 1, 2);
    else PLBM_ADJ(0, 1, 2);
  } else {
    if (field) PLBM_ADJ(2, 0, 2);
    else if (uniform) PLBM_ADJ(1, 0, 2);
    else PLBM_ADJ(0, 0, 2);
  }
#undef PLBM_ADJ
  return (int)hipGetLastError();
}

// part == nullptr: no rate sensitivities (rollouts that do not differentiate tau skip
// their per-population sums and the workgroup reduction)
template <class L, int COLL, int ST, class LT, int PAT>
int adj_step_tp(const KArgs& a, bool field, const void* f, const void* lin, void* lout,
                float* dF, double* part, hipStream_t s) {
  if (part) return adj_step_tpr<L, COLL, ST, LT, PAT, true>(a, field, f, lin, lout, dF, part, s);
  return adj_step_tpr<L, COLL, ST, LT, PAT, false>(a, field, f, lin, lout, dF, part, s);
}

template <class L, int COLL, int ST, class LT>
int adj_step_t(const KArgs& a, int pattern, bool field, const void* f, const void* lin,
               void* lout, float* dF, double* part, hipStream_t s) {
  switch (pattern) {
    case PYLBM_AB: return adj_step_tp<L, COLL, ST, LT, PYLBM_AB>(a, field, f, lin, lout, dF, part, s);
    case PYLBM_AA_EVEN:
      return adj_step_tp<L, COLL, ST, LT, PYLBM_AA_EVEN>(a, field, f, lin, lout, dF, part, s);
    case PYLBM_AA_ODD:
      return adj_step_tp<L, COLL, ST, LT, PYLBM_AA_ODD>(a, field, f, lin, lout, dF, part, s);
  }
  return PYLBM_EINVAL;
}

template <class L, int COLL, class LT>
int adj_step_s(const KArgs& a, int storage, int pattern, bool field, const void* f,
               const void* lin, void* lout, float* dF, double* part, hipStream_t s) {
  switch (storage) {
    case PYLBM_F32:
      return adj_step_t<L, COLL, PYLBM_F32, LT>(a, pattern, field, f, lin, lout, dF, part, s);
    case PYLBM_F16:
      return adj_step_t<L, COLL, PYLBM_F16, LT>(a, pattern, field, f, lin, lout, dF, part, s);
    case PYLBM_BF16:
      return adj_step_t<L, COLL, PYLBM_BF16, LT>(a, pattern, field, f, lin, lout, dF, part, s);
    case PYLBM_F32S:
      return adj_step_t<L, COLL, PYLBM_F32S, LT>(a, pattern, field, f, lin, lout, dF, part, s);
  }
  return PYLBM_EINVAL;
}

template <class L, class LT>
int adj_step_l(const KArgs& a, int coll, int storage, int pattern, bool field, const void* f,
               const void* lin, void* lout, float* dF, double* part, hipStream_t s) {
  if (coll == PYLBM_BGK)
    return adj_step_s<L, PYLBM_BGK, LT>(a, storage, pattern, field, f, lin, lout, dF, part, s);
  if (coll == PYLBM_TRT)
    return adj_step_s<L, PYLBM_TRT, LT>(a, storage, pattern, field, f, lin, lout, dF, part, s);
  return PYLBM_ENOTSUP;
}


// Record mode: one fp32 record array of N x 16 bytes in place of the stored state; the
// arrays are under 4 GiB (the host checks), so one descriptor serves each.
template <class L, int COLL, class LT, int PAT>
int adj_rec_tp(const KArgs& a, bool field, const void* rec, const void* lin, void* lout,
               float* dF, hipStream_t s) {
  const dim3 grid((unsigned)a.nblk), block(kBlock);
  const bool uniform = a.fx != 0.0f || a.fy != 0.0f || a.fz != 0.0f;
  const float* r = (const float*)rec;
  const LT* li = (const LT*)lin;
  LT* lo = (LT*)lout;
#define PLBM_ADJR(FMv, SPv)                                                                  \
  k_adjoint<L, COLL, PYLBM_F32, FMv, 1, LT, SPv, PAT, false, true><<<grid, block, 0, s>>>( \
      a, r, li, lo, dF, nullptr)
  if (a.resident) {
    if (field) PLBM_ADJR(2, 0);
    else if (uniform) PLBM_ADJR(1, 0);
    else PLBM_ADJR(0, 0);
  } else {
    if (field) PLBM_ADJR(2, 2);
    else if (uniform) PLBM_ADJR(1, 2);
    else PLBM_ADJR(0, 2);
  }
#undef PLBM_ADJR
  return (int)hipGetLastError();
}

template <class L, int COLL, class LT>
int adj_rec_t(const KArgs& a, int pattern, bool field, const void* rec, const void* lin,
              void* lout, float* dF, hipStream_t s) {
  switch (pattern) {
    case PYLBM_AB: return adj_rec_tp<L, COLL, LT, PYLBM_AB>(a, field, rec, lin, lout, dF, s);
    case PYLBM_AA_EVEN:
      return adj_rec_tp<L, COLL, LT, PYLBM_AA_EVEN>(a, field, rec, lin, lout, dF, s);
    case PYLBM_AA_ODD:
      return adj_rec_tp<L, COLL, LT, PYLBM_AA_ODD>(a, field, rec, lin, lout, dF, s);
  }
  return PYLBM_EINVAL;
}

// One adjoint step of lattice L with adjoint states of type LT (each (L, LT) pair is
// instantiated in its own translation unit, adj_<lattice>_<type>.hip, so that the four
// compile in parallel).  rec: f is a moment record (part must be nullptr).
template <class L, class LT>
int adj_step_entry(const KArgs& a, int coll, int storage, int pattern, bool field, bool rec,
                   const void* f, const void* lin, void* lout, float* dF, double* part,
                   hipStream_t s) {
  if (rec) {
    if (part) return PYLBM_EINVAL;
    if (coll == PYLBM_BGK) return adj_rec_t<L, PYLBM_BGK, LT>(a, pattern, field, f, lin, lout, dF, s);
    if (coll == PYLBM_TRT) return adj_rec_t<L, PYLBM_TRT, LT>(a, pattern, field, f, lin, lout, dF, s);
    return PYLBM_ENOTSUP;
  }
  return adj_step_l<L, LT>(a, coll, storage, pattern, field, f, lin, lout, dF, part, s);
}


}  // namespace adj
}  // namespace plbm
