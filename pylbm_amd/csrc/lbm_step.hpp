// The fused stream-collide kernel for gfx950, templated over lattice, collision, storage,
// streaming pattern and force mode.
//
// Semantics (what is read and written per node) are those of the reference's fused kernels
// (pylbm/codegen.py:10-16 for the patterns, pylbm/backends/cuda/lbm.cu:71-98 and 165-198
// for pull / bounce-back / Ladd / AA push); the structure is this file's own (memory access
// in lbm_access.hpp): one lane per node, scalar-unit row logic, shared 32-bit lane offsets,
// raw buffer access, a branch-free gather, the collision on deviations f - w
// (lbm_common.hpp), non-temporal stores for the two-buffer and AA-even patterns, and an
// XCD-aware workgroup order.  One node per lane with natural alignment measured faster on
// MI355X than 2, 4 or 8 nodes per lane with unaligned 8/16-byte accesses along x.
#pragma once

#include "lbm_access.hpp"

namespace plbm {

enum ForceMode { kForceNone = 0, kForceUniform = 1, kForceField = 2 };

template <class L, int COLL, int ST, int PAT, int FM, bool SMALL>
__global__ void __launch_bounds__(kBlock)
k_step(const KArgs a, const typename Codec<ST>::T* A, typename Codec<ST>::T* B) {
  using C = Codec<ST>;
  using T = typename C::T;
  constexpr int Q = L::Q;
  constexpr uint32_t TS = sizeof(T);

  Lane l;
  if (!lane_setup<TS>(a, remap_block((int)blockIdx.x, a.nblk, a.remap), l)) return;
  const size_t N = (size_t)a.N;
  const char* Ab = reinterpret_cast<const char*>(A);
  char* Bb = reinterpret_cast<char*>(B);
  const size_t plane = N * TS;

  float h[Q];  // deviations f - w of the pulled populations
  gather<L, ST, PAT, SMALL>(a, l, Ab, plane, h);

  float F[3] = {0.f, 0.f, 0.f};
  if (FM == kForceUniform) {
    F[0] = a.fx;
    F[1] = a.fy;
    F[2] = a.fz;
  } else if (FM == kForceField) {
    const char* ffb = reinterpret_cast<const char*>(a.ff);
    const size_t fplane = N * sizeof(float);
    const uint32_t of = l.n * 4u;
    F[0] = ld_at<float>(ffb, of);
    F[1] = ld_at<float>(ffb + fplane, of);
    F[2] = (L::D == 3) ? ld_at<float>(ffb + 2 * fplane, of) : 0.0f;
  }
  float d[Q];
  collide_node<L, COLL, FM != kForceNone>(h, d, F, a.r0, a.r1, a.r2, a.r3);

#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const float v = C::shifted ? d[i] : d[i] + (float)L::w(i);
    if (PAT == PYLBM_AB) {
      st_pl<2, SMALL>(Bb, plane, i, 0, l.o0, C::st(v));
      continue;
    }
    if (PAT == PYLBM_AA_EVEN) {
      st_pl<2, SMALL>(Bb, plane, L::opp(i), 0, l.o0, C::st(v));
      continue;
    }
    if (i == 0) {
      st_pl<0, SMALL>(Bb, plane, 0, 0, l.o0, C::st(v));
      continue;
    }
    // AA odd: push to x + c_i, or bounce back into the own slot opp(i)
    const int cx = L::c(i, 0);
    const RowLink r = row_link<L>(a, l.y, l.z, i, +1);
    if (r.wall) {
      float vb = v;
      if (cx != 0 && r.top) vb = v + (float)(-6.0 * L::w(i) * cx) * a.ulid;
      st_pl<0, SMALL>(Bb, plane, L::opp(i), 0, l.o0, C::st(vb));
      continue;
    }
    const uint32_t off = cx > 0 ? l.op : (cx < 0 ? l.om : l.o0);  // push target x + c_x
    if ((cx > 0 && l.wall_hi) || (cx < 0 && l.wall_lo))
      st_pl<0, SMALL>(Bb, plane, L::opp(i), 0, l.o0, C::st(v));
    else
      st_pl<0, SMALL>(Bb, plane, i, r.delta * (ptrdiff_t)TS, off, C::st(v));
  }
}

template <class L, int COLL, int ST, int PAT, bool SMALL>
void launch_step_f(const KArgs& a, int fmode, const void* src, void* dst, hipStream_t s) {
  using T = typename Codec<ST>::T;
  const dim3 grid((unsigned)a.nblk), block(kBlock);
  if (fmode == kForceField)
    k_step<L, COLL, ST, PAT, kForceField, SMALL><<<grid, block, 0, s>>>(a, (const T*)src,
                                                                        (T*)dst);
  else if (fmode == kForceUniform)
    k_step<L, COLL, ST, PAT, kForceUniform, SMALL><<<grid, block, 0, s>>>(a, (const T*)src,
                                                                          (T*)dst);
  else
    k_step<L, COLL, ST, PAT, kForceNone, SMALL><<<grid, block, 0, s>>>(a, (const T*)src,
                                                                       (T*)dst);
}

template <class L, int COLL, int ST, int PAT>
int launch_step_t(const KArgs& a, int fmode, const void* src, void* dst, hipStream_t s) {
  // one descriptor per array when the population array and the force field fit in 4 GiB
  const uint64_t bytes = (uint64_t)L::Q * (uint64_t)a.N * sizeof(typename Codec<ST>::T);
  const uint64_t fbytes = (uint64_t)L::D * (uint64_t)a.N * sizeof(float);
  if (bytes < ((uint64_t)1 << 32) && fbytes < ((uint64_t)1 << 32))
    launch_step_f<L, COLL, ST, PAT, true>(a, fmode, src, dst, s);
  else
    launch_step_f<L, COLL, ST, PAT, false>(a, fmode, src, dst, s);
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
