// The fused stream-collide kernel for gfx950, templated over lattice, collision, storage,
// streaming pattern and force mode.
//
// Semantics (what is read and written per node) are those of the reference's fused kernels
// (pylbm/codegen.py:10-16 for the patterns, pylbm/backends/cuda/lbm.cu:71-98 and 165-198
// for pull / bounce-back / Ladd / AA push); the structure is this file's own (memory access
// in lbm_access.hpp): one lane per node, scalar-unit row logic, shared 32-bit lane offsets,
// raw buffer access, a branch-free gather, the collision on deviations f - w
// (lbm_common.hpp), non-temporal stores for the two-buffer and AA-even patterns, and an
// XCD-aware workgroup order.  A lane serves 1, 2 or 4 nodes one workgroup width apart
// (nodes_per_lane), so every access stays naturally aligned; 2, 4 or 8 adjacent nodes per
// lane with unaligned 8/16-byte accesses along x measured slower.
#pragma once

#include "lbm_access.hpp"

namespace plbm {

enum ForceMode { kForceNone = 0, kForceUniform = 1, kForceField = 2 };

// Store the post-collision deviations d of one node for pattern PAT
// (pylbm/codegen.py:10-16): AB -> B[i, x]; AA even -> A[opp(i), x]; AA odd -> A[i, x + c_i],
// or the own slot opp(i) on a wall link (minus the lid term through the lid).
// SP: cache policy of the two-buffer and AA-even stores (2 non-temporal for states larger
// than the Infinity Cache, 0 for resident states, whose next launch re-reads them).
template <class L, int ST, int PAT, int ADDR, int SP>
__device__ __forceinline__ void store_node(const KArgs& a, const Lane& l, char* Bb,
                                           size_t plane, const float (&d)[L::Q]) {
  using C = Codec<ST>;
  using T = typename C::T;
  constexpr int Q = L::Q;
  constexpr uint32_t TS = sizeof(T);
  // AA odd, waves with no lane on an x wall (every wave of a box periodic in x): each push
  // goes to x + c_i or, through a y/z wall (uniform, selected by value), to the own slot,
  // without the per-lane branches (each a divergent branch with exec-mask bookkeeping)
  if (PAT == PYLBM_AA_ODD && __ballot(l.wall_lo || l.wall_hi) == 0) {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const float v = C::shifted ? d[i] : d[i] + (float)L::w(i);
      const int cx = L::c(i, 0);
      const RowLink r = row_link<L>(a, l.y, l.z, i, +1);
      const bool own = i == 0 || r.wall;
      const float lid = (cx != 0 && r.wall && r.top) ? (float)(-6.0 * L::w(i) * cx) * a.ulid
                                                    : 0.0f;
      st_pl<0, ADDR>(Bb, plane, own ? L::opp(i) : i, own ? 0 : r.delta * (ptrdiff_t)TS,
                     own ? l.o0 : (cx > 0 ? l.op : (cx < 0 ? l.om : l.o0)), C::st(v + lid));
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const float v = C::shifted ? d[i] : d[i] + (float)L::w(i);
    if (PAT == PYLBM_AB) {
      st_pl<SP, ADDR>(Bb, plane, i, 0, l.o0, C::st(v));
      continue;
    }
    if (PAT == PYLBM_AA_EVEN) {
      st_pl<SP, ADDR>(Bb, plane, L::opp(i), 0, l.o0, C::st(v));
      continue;
    }
    if (i == 0) {
      st_pl<0, ADDR>(Bb, plane, 0, 0, l.o0, C::st(v));
      continue;
    }
    // AA odd: push to x + c_i, or bounce back into the own slot opp(i)
    const int cx = L::c(i, 0);
    const RowLink r = row_link<L>(a, l.y, l.z, i, +1);
    if (r.wall) {
      float vb = v;
      if (cx != 0 && r.top) vb = v + (float)(-6.0 * L::w(i) * cx) * a.ulid;
      st_pl<0, ADDR>(Bb, plane, L::opp(i), 0, l.o0, C::st(vb));
      continue;
    }
    const uint32_t off = cx > 0 ? l.op : (cx < 0 ? l.om : l.o0);  // push target x + c_x
    if ((cx > 0 && l.wall_hi) || (cx < 0 && l.wall_lo))
      st_pl<0, ADDR>(Bb, plane, L::opp(i), 0, l.o0, C::st(v));
    else
      st_pl<0, ADDR>(Bb, plane, i, r.delta * (ptrdiff_t)TS, off, C::st(v));
  }
}

// NPL nodes per lane: lane t of workgroup (row, seg) serves x = seg * kBlock * NPL +
// k * kBlock + t for k < NPL, so every wave access stays one contiguous, aligned run of 64
// elements; the loads of all NPL nodes are issued before the first collision, which
// doubles the bytes a wave keeps in flight when NPL = 2.
// REC: also write the node's moment record (rho, u as the collision used them) to a.rec
// (non-temporal: the records of a segment are read once, by its adjoint steps).
template <class L, int COLL, int ST, int PAT, int FM, int ADDR, int NPL, int SP,
          bool REC = false>
__global__ void __launch_bounds__(kBlock)
k_step(const KArgs a, const typename Codec<ST>::T* A, typename Codec<ST>::T* B) {
  using C = Codec<ST>;
  using T = typename C::T;
  constexpr int Q = L::Q;
  constexpr uint32_t TS = sizeof(T);

  int y, z, seg;
  if (!block_row(a, remap_block((int)blockIdx.x, a), y, z, seg)) return;
  Lane l[NPL];
  bool act[NPL];
#pragma unroll
  for (int k = 0; k < NPL; ++k)
    act[k] = lane_at<TS>(a, y, z, seg * (kBlock * NPL) + k * kBlock + seg_lane(a), l[k]);
  if (!act[0]) return;
#pragma unroll
  for (int k = 1; k < NPL; ++k)
    if (!act[k]) l[k] = l[0];  // past the row end: re-read node 0, store nothing
  const size_t N = (size_t)a.N;
  const char* Ab = reinterpret_cast<const char*>(A);
  char* Bb = reinterpret_cast<char*>(B);
  const size_t plane = N * TS;
  // non-temporal loads of the x-unshifted links for the streamed two-buffer pattern
  // (A/B on MI355X, 4096^2 BGK: fp32 +1.2%, fp16 +0.6%; D3Q19 512^3 TRT unchanged; all
  // links non-temporal: -1% to -9%; AA patterns: -3%)
  constexpr int LA = (SP == 2 && PAT == PYLBM_AB) ? 2 : 0;

  T t[NPL][Q];
  float lid[NPL][Q];
  float F[NPL][3];
#pragma unroll
  for (int k = 0; k < NPL; ++k) {
    gather_issue<L, ST, PAT, ADDR, LA, 0>(a, l[k], Ab, plane, t[k], lid[k]);
    F[k][0] = F[k][1] = F[k][2] = 0.0f;
    if (FM == kForceUniform) {
      F[k][0] = a.fx;
      F[k][1] = a.fy;
      F[k][2] = a.fz;
    } else if (FM == kForceField) {  // issued with the gather, before any wait
      const char* ffb = reinterpret_cast<const char*>(a.ff);
      const size_t fplane = N * sizeof(float);
      const uint32_t of = l[k].n * 4u;
      F[k][0] = ld_at<float>(ffb, of);
      F[k][1] = ld_at<float>(ffb + fplane, of);
      if (L::D == 3) F[k][2] = ld_at<float>(ffb + 2 * fplane, of);
    }
  }

#pragma unroll
  for (int k = 0; k < NPL; ++k) {
    float h[Q];  // deviations f - w of the pulled populations
    gather_finish<L, ST, PAT, ADDR>(l[k], Ab, plane, t[k], lid[k], h);
    float d[Q], mo[4];
    collide_node<L, COLL, FM != kForceNone>(h, d, F[k], a.r0, a.r1, a.r2, a.r3, mo);
    if (k > 0 && !act[k]) break;
    store_node<L, ST, PAT, ADDR, SP>(a, l[k], Bb, plane, d);
    if constexpr (REC) st_rec<2>(a.rec, l[k].n * 16u, mo);
  }
}

// Nodes per lane, from A/B timing on MI355X (same box, alternating runs, with the
// workgroup order of set_xcd_order): serving more nodes per lane keeps more bytes in
// flight per wave, which pays until register pressure cuts occupancy or the longer row
// segments spoil the DRAM locality of the common row window.  D2Q9: one node per lane
// for the two-buffer kernels (4096^2 fp16 +5%, bf16 +4%, fp32 +3.6% over two; fp16 165k
// vs 157k MLUPS), two for the 16-bit AA kernels (+1-2% over one), one for fp32 AA.
// D3Q19 (19 populations per node, 512^3 TRT): two for 16-bit (bf16 68.8k vs 66.0k MLUPS at
// one, fp16 and AA-fp16 within 0.5%), one for fp32 (two loses 3%).
template <class L, int ST, int PAT>
constexpr int nodes_per_lane() {
  constexpr bool half = sizeof(typename Codec<ST>::T) == 2;
  if constexpr (L::Q == 9) return (half && PAT != PYLBM_AB) ? 2 : 1;
  else return half ? 2 : 1;
}

// The row decomposition for NPL nodes per lane (segments of kBlock * NPL nodes; only used
// on rows of at least kBlock * NPL nodes, which one workgroup per segment serves).
inline KArgs with_npl(const KArgs& a, int npl) {
  if (npl == 1) return a;
  KArgs b = a;
  b.sw_log2 = 8;
  b.rpw_log2 = 0;
  b.nseg = (a.nx + kBlock * npl - 1) / (kBlock * npl);
  b.nblk = b.nseg * a.ny * a.nz;
  set_xcd_order(b, a.order, npl);
  b.div_nseg = make_fastdiv((uint32_t)b.nseg);
  return b;
}

template <class L, int COLL, int ST, int PAT, int ADDR, int NPL, int SP, bool REC = false>
void launch_step_n(const KArgs& a0, int fmode, const void* src, void* dst, hipStream_t s) {
  using T = typename Codec<ST>::T;
  const KArgs a = with_npl(a0, NPL);
  const dim3 grid((unsigned)a.nblk), block(kBlock);
  const T* in = (const T*)src;
  T* out = (T*)dst;
  if (fmode == kForceField)
    k_step<L, COLL, ST, PAT, kForceField, ADDR, NPL, SP, REC><<<grid, block, 0, s>>>(a, in, out);
  else if (fmode == kForceUniform)
    k_step<L, COLL, ST, PAT, kForceUniform, ADDR, NPL, SP, REC><<<grid, block, 0, s>>>(a, in,
                                                                                      out);
  else
    k_step<L, COLL, ST, PAT, kForceNone, ADDR, NPL, SP, REC><<<grid, block, 0, s>>>(a, in, out);
}

// Several nodes per lane only pay on rows at least one full segment long; shorter rows
// would leave the extra node slots idle and keep their register cost.
template <class L, int COLL, int ST, int PAT, int ADDR, int SP>
void launch_step_f(const KArgs& a, int fmode, const void* src, void* dst, hipStream_t s) {
  constexpr int NPL = nodes_per_lane<L, ST, PAT>();
  if constexpr (NPL > 1) {
    if (a.nx >= kBlock * NPL) {
      launch_step_n<L, COLL, ST, PAT, ADDR, NPL, SP>(a, fmode, src, dst, s);
      return;
    }
  }
  launch_step_n<L, COLL, ST, PAT, ADDR, 1, SP>(a, fmode, src, dst, s);
}

// Addressing mode of a state array of q planes (plane_ref in lbm_access.hpp): one
// descriptor below 4 GiB; above, plane groups of 8 or 4 slots when (G + 2) planes fit in
// 4 GiB (D3Q19 512^3: fp16 planes of 256 MiB -> 8, fp32 planes of 512 MiB -> 4), else one
// descriptor per access.  Few descriptors keep the uniform addressing in scalar registers:
// with one per plane and row, the two-nodes-per-lane D3Q19 kernels spill scalar registers
// into vector lanes (~120 vector instructions per node).  Same-box A/B at 512^3 against
// per-access descriptors: AA fp16 +4.5%, two-buffer fp16 +1.3%, fp32 +0.5%, bf16 -0.2%.
inline int addr_mode(uint64_t q, uint64_t plane) {
  const uint64_t lim = (uint64_t)1 << 32;
  if (q * plane < lim) return 1;
  if (10 * plane <= lim) return 8;
  if (6 * plane <= lim) return 4;
  return 0;
}

template <class L, int COLL, int ST, int PAT>
int launch_step_t(const KArgs& a, int fmode, const void* src, void* dst, hipStream_t s) {
  const uint64_t plane = (uint64_t)a.N * sizeof(typename Codec<ST>::T);
  const uint64_t fbytes = (uint64_t)L::D * (uint64_t)a.N * sizeof(float);
  const int mode = fbytes < ((uint64_t)1 << 32) ? addr_mode(L::Q, plane) : 0;
  if (a.rec) {
    // moment records: the two-buffer pattern of the BGK / TRT rollouts, arrays under 4 GiB,
    // one node per lane (the host checks the sizes)
    if constexpr (PAT == PYLBM_AB && COLL != PYLBM_MRT) {
      if (mode != 1) return PYLBM_ERANGE;
      if (a.resident) launch_step_n<L, COLL, ST, PAT, 1, 1, 0, true>(a, fmode, src, dst, s);
      else launch_step_n<L, COLL, ST, PAT, 1, 1, 2, true>(a, fmode, src, dst, s);
      return (int)hipGetLastError();
    }
    return PYLBM_ENOTSUP;
  }
  // a state resident in the Infinity Cache (always under 4 GiB) is stored with temporal
  // stores, so that the next launch finds it there; larger states stream past it
  switch (mode) {
    case 1:
      if (a.resident) launch_step_f<L, COLL, ST, PAT, 1, 0>(a, fmode, src, dst, s);
      else launch_step_f<L, COLL, ST, PAT, 1, 2>(a, fmode, src, dst, s);
      break;
    case 8: launch_step_f<L, COLL, ST, PAT, 8, 2>(a, fmode, src, dst, s); break;
    case 4: launch_step_f<L, COLL, ST, PAT, 4, 2>(a, fmode, src, dst, s); break;
    default: launch_step_f<L, COLL, ST, PAT, 0, 2>(a, fmode, src, dst, s); break;
  }
  return (int)hipGetLastError();
}

template <class L, int COLL, int ST>
int launch_step_p(const KArgs& a, int pattern, int fmode, const void* src, void* dst,
                  hipStream_t s) {
  switch (pattern) {
    case PYLBM_AB: return launch_step_t<L, COLL, ST, PYLBM_AB>(a, fmode, src, dst, s);
    case PYLBM_AA_ODD: return launch_step_t<L, COLL, ST, PYLBM_AA_ODD>(a, fmode, src, dst, s);
    case PYLBM_AA_EVEN: return launch_step_t<L, COLL, ST, PYLBM_AA_EVEN>(a, fmode, src, dst, s);
  }
  return PYLBM_EINVAL;
}

template <class L, int COLL>
int launch_step_s(const KArgs& a, int storage, int pattern, int fmode, const void* src,
                  void* dst, hipStream_t s) {
  switch (storage) {
    case PYLBM_F32: return launch_step_p<L, COLL, PYLBM_F32>(a, pattern, fmode, src, dst, s);
    case PYLBM_F16: return launch_step_p<L, COLL, PYLBM_F16>(a, pattern, fmode, src, dst, s);
    case PYLBM_BF16: return launch_step_p<L, COLL, PYLBM_BF16>(a, pattern, fmode, src, dst, s);
    case PYLBM_F32S: return launch_step_p<L, COLL, PYLBM_F32S>(a, pattern, fmode, src, dst, s);
  }
  return PYLBM_EINVAL;
}

// Host-side dispatcher for one lattice; each lattice is instantiated in its own
// translation unit (step_d2q9.hip, step_d3q19.hip) so they compile in parallel.
template <class L>
int launch_step(const KArgs& a, int coll, int storage, int pattern, int fmode,
                const void* src, void* dst, hipStream_t s) {
  switch (coll) {
    case PYLBM_BGK: return launch_step_s<L, PYLBM_BGK>(a, storage, pattern, fmode, src, dst, s);
    case PYLBM_TRT: return launch_step_s<L, PYLBM_TRT>(a, storage, pattern, fmode, src, dst, s);
    case PYLBM_MRT:
      if constexpr (L::Q == 9)
        return launch_step_s<L, PYLBM_MRT>(a, storage, pattern, fmode, src, dst, s);
      else
        return PYLBM_ENOTSUP;
  }
  return PYLBM_EINVAL;
}

}  // namespace plbm
