// General geometry for the two-buffer fused step: stationary solid nodes, the velocity
// inlet on the x = 0 wall and the rho = 1 pressure outlet on the last column
// (pylbm/core.py:56-148; the reference runs these only in its unfused NumPy / torch
// solvers, pylbm/backends/torch_backend.py:80-137, with a q*N int64 gather table).
//
// Here the geometry costs one extra load per node: a bit mask of the directions whose pull
// source x - c_i is solid (uint16 for D2Q9, uint32 for D3Q19; built once by k_links).  The
// gather is the box gather of lbm_access.hpp (row logic on the scalar unit, all loads
// issued back to back); lanes whose mask is nonzero, or that sit on an x wall, then read
// the own opp(i) slots in one divergent block.  The value of a link is
//   own opp(i) slot          if the source is solid or behind any wall, else the pulled one,
//   + Ladd lid term          if the last walled axis it crosses is the lid,
//   + 6 w_i u_in(y)          if the last walled axis it crosses is the x = 0 inlet wall,
// which is the reference's order of application (solid, then walls axis by axis, later
// axes overriding earlier ones).  The outlet then replaces the c_x = -1 populations of
// the last column by the rho = 1 equilibrium at the velocity of the stored state one
// column to the left.
#pragma once

#include <type_traits>

#include "lbm_step.hpp"

namespace plbm {

template <class L>
using LinkMask = typename std::conditional<(L::Q > 16), uint32_t, uint16_t>::type;

// bit i set for the directions with c_x > 0 (resp. < 0)
template <class L>
__host__ __device__ constexpr uint32_t x_bits(int sign) {
  uint32_t m = 0;
  for (int i = 1; i < L::Q; ++i)
    if (L::c(i, 0) * sign > 0) m |= 1u << i;
  return m;
}

// Group flag of the wave that serves node slot l: the wave covers 64 consecutive nodes of
// one row, aligned to 64 (x of its first active lane, over 64, names the group), so the
// flag is uniform and read once per wave by a scalar load.
__device__ __forceinline__ bool group_flag(const KArgs& a, const Lane& l) {
  const int row = l.y + a.ny * l.z;
  const int xg = __builtin_amdgcn_readfirstlane(l.x) >> 6;
  return a.lgroups[row * a.gpr + xg] != 0u;
}

// Loads of one node: the pulled values (box rules), their additive terms, the mask.
// GF (group flags): gf says whether any lane of the wave has a solid source or an x wall;
// if so, the own opp(i) slots of every lane are loaded here too (tw), with the pull loads,
// and the choice between them waits for the mask; if not, neither the mask nor the own
// slots are loaded.  Without GF the mask is loaded and the own slots follow it (geo_finish).
template <class L, int ST, bool GF>
__device__ __forceinline__ void geo_issue(const KArgs& a, const Lane& l, const char* Ab,
                                          size_t plane, float uin, bool gf,
                                          typename Codec<ST>::T (&t)[L::Q], float (&add)[L::Q],
                                          typename Codec<ST>::T (&tw)[L::Q], uint32_t& bb) {
  using T = typename Codec<ST>::T;
  using M = LinkMask<L>;
  constexpr int Q = L::Q;
  constexpr uint32_t TS = sizeof(T);
  bb = 0;
  // without group flags the mask goes first: the own-slot loads of geo_finish wait for it
  if (!GF && a.links)
    bb = (uint32_t)ld_at<M>(reinterpret_cast<const char*>(a.links), l.n * sizeof(M));
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int cx = L::c(i, 0);
    add[i] = 0.0f;
    if (i == 0) {
      t[0] = ld_pl<T, 1>(Ab, plane, 0, 0, l.o0);
      continue;
    }
    const RowLink r = row_link<L>(a, l.y, l.z, i, -1);
    const uint32_t offx = cx > 0 ? l.om : (cx < 0 ? l.op : l.o0);
    t[i] = ld_pl<T, 1>(Ab, plane, r.wall ? L::opp(i) : i,
                          r.wall ? 0 : r.delta * (ptrdiff_t)TS, r.wall ? l.o0 : offx);
    if (cx != 0 && r.wall && r.top) add[i] = (float)(6.0 * L::w(i) * cx) * a.ulid;
    // inlet: crosses the x = 0 wall and no later (y / z) wall
    if (L::D == 2 && cx == 1 && !r.wall)
      add[i] = l.wall_lo ? (float)(6.0 * L::w(i)) * uin : 0.0f;
  }
  // with group flags the pull loads go first: they do not wait for the scalar flag load
  if (GF && gf) {
    bb = (uint32_t)ld_at<M>(reinterpret_cast<const char*>(a.links), l.n * sizeof(M));
#pragma unroll
    for (int i = 1; i < Q; ++i) tw[i] = ld_pl<T, 1>(Ab, plane, L::opp(i), 0, l.o0);
  }
}

// Bounce-back of solid sources and x walls, decode, outlet: the deviations h of one node.
template <class L, int ST, bool GF>
__device__ __forceinline__ void geo_finish(const KArgs& a, const Lane& l, const char* Ab,
                                           size_t plane, typename Codec<ST>::T (&t)[L::Q],
                                           const float (&add)[L::Q],
                                           typename Codec<ST>::T (&tw)[L::Q], uint32_t bb,
                                           float (&h)[L::Q]) {
  using C = Codec<ST>;
  using T = typename C::T;
  constexpr int Q = L::Q;
  if (l.wall_lo) bb |= x_bits<L>(+1);
  if (l.wall_hi) bb |= x_bits<L>(-1);
  // with GF a lane with bb != 0 belongs to a flagged group, whose own slots are loaded
  if (bb != 0) {
    if (!GF) {
#pragma unroll
      for (int i = 1; i < Q; ++i) tw[i] = ld_pl<T, 1>(Ab, plane, L::opp(i), 0, l.o0);
    }
#pragma unroll
    for (int i = 1; i < Q; ++i)
      if ((bb >> i) & 1u) t[i] = tw[i];
  }
#pragma unroll
  for (int i = 0; i < Q; ++i) h[i] = C::ld(t[i]) + (C::shifted ? 0.0f : -(float)L::w(i)) + add[i];

  if (L::D == 2 && a.outlet && l.x == a.nx - 1) {
    // rho = 1 equilibrium at the velocity of the stored column x - 1 (no force correction)
    T tc[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) tc[i] = ld_pl<T, 1>(Ab, plane, i, 0, l.om);
    float rho = 1.0f, jx = 0.0f, jy = 0.0f;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const float g = C::ld(tc[i]) + (C::shifted ? 0.0f : -(float)L::w(i));
      rho += g;
      jx += (float)L::c(i, 0) * g;
      jy += (float)L::c(i, 1) * g;
    }
    const float ux = jx / rho, uy = jy / rho;
    const float usq = ux * ux + uy * uy;
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      if (L::c(i, 0) != -1) continue;
      const float cu = -ux + (float)L::c(i, 1) * uy;
      h[i] = (float)L::w(i) * (3.0f * cu + 4.5f * cu * cu - 1.5f * usq);
    }
  }
}

// NPL nodes per lane, one workgroup width apart, as k_step (lbm_step.hpp).  GF: the group
// flags of a.lgroups select, per wave, the one-round gather of solid waves (geo_issue).
template <class L, int COLL, int ST, int FM, int NPL, bool GF>
__global__ void __launch_bounds__(kBlock)
k_step_geo(const KArgs a, const typename Codec<ST>::T* A, typename Codec<ST>::T* B) {
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
  const size_t plane = (size_t)a.N * TS;
  const char* Ab = reinterpret_cast<const char*>(A);
  char* Bb = reinterpret_cast<char*>(B);
  const float uin = a.inlet ? a.inlet[y] : 0.0f;  // uniform: one row per workgroup

  T t[NPL][Q], tw[NPL][Q];
  float add[NPL][Q];
  uint32_t bb[NPL];
  float F[NPL][3];
  bool gf[NPL];
#pragma unroll
  for (int k = 0; k < NPL; ++k) gf[k] = GF ? group_flag(a, l[k]) : true;
#pragma unroll
  for (int k = 0; k < NPL; ++k) {
    geo_issue<L, ST, GF>(a, l[k], Ab, plane, uin, gf[k], t[k], add[k], tw[k], bb[k]);
    F[k][0] = F[k][1] = F[k][2] = 0.0f;
    if (FM == 1) {
      F[k][0] = a.fx;
      F[k][1] = a.fy;
      F[k][2] = a.fz;
    } else if (FM == 2) {
      const char* ffb = reinterpret_cast<const char*>(a.ff);
      const size_t fplane = (size_t)a.N * sizeof(float);
      const uint32_t of = l[k].n * 4u;
      F[k][0] = ld_at<float>(ffb, of);
      F[k][1] = ld_at<float>(ffb + fplane, of);
      if (L::D == 3) F[k][2] = ld_at<float>(ffb + 2 * fplane, of);
    }
  }

#pragma unroll
  for (int k = 0; k < NPL; ++k) {
    float h[Q];
    geo_finish<L, ST, GF>(a, l[k], Ab, plane, t[k], add[k], tw[k], bb[k], h);
    float d[Q];
    collide_node<L, COLL, FM != 0>(h, d, F[k], a.r0, a.r1, a.r2, a.r3);
    if (k > 0 && !act[k]) break;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const float v = C::shifted ? d[i] : d[i] + (float)L::w(i);
      st_pl<2, 1>(Bb, plane, i, 0, l[k].o0, C::st(v));
    }
  }
}

// Per-node solid-source mask: bit i (i >= 1) set when the node x - c_i, wrapped
// periodically on every axis as np.roll does (pylbm/core.py:66-71), is solid.
template <class L>
__global__ void __launch_bounds__(256)
k_links(const uint8_t* solid, LinkMask<L>* links, int nx, int ny, int nz, int N) {
  const int n = (int)(blockIdx.x * 256 + threadIdx.x);
  if (n >= N) return;
  const int x = n % nx, r = n / nx, y = r % ny, z = r / ny;
  uint32_t m = 0;
#pragma unroll
  for (int i = 1; i < L::Q; ++i) {
    const int sx = (x - L::c(i, 0) + nx) % nx;
    const int sy = (y - L::c(i, 1) + ny) % ny;
    const int sz = (z - L::c(i, 2) + nz) % nz;
    if (solid[sx + nx * (sy + ny * sz)]) m |= 1u << i;
  }
  links[n] = (LinkMask<L>)m;
}

// Nodes per lane of the geometry kernel (D2Q9 only on rows at least one segment long),
// from A/B timing on the cylinder-array channel, 4096^2 TRT: one for fp32 (80.5k MLUPS vs
// 76.8k at two), four for 16-bit storage (fp16 139.9k vs 138.3k at one, bf16 138.2k vs
// 133.0k); the mask load and the bounce-back block shift the optimum away from the box
// kernel's.  One for D3Q19, where the bounce-back block's extra registers make two nodes
// per lane 6% slower (sphere-array box, 256^3, fp16).
template <class L, int COLL, int ST, int NPL, bool GF>
void launch_geo_g(const KArgs& a, int fmode, const void* src, void* dst, hipStream_t s) {
  using T = typename Codec<ST>::T;
  const dim3 grid((unsigned)a.nblk), block(kBlock);
  if (fmode == 2)
    k_step_geo<L, COLL, ST, 2, NPL, GF><<<grid, block, 0, s>>>(a, (const T*)src, (T*)dst);
  else if (fmode == 1)
    k_step_geo<L, COLL, ST, 1, NPL, GF><<<grid, block, 0, s>>>(a, (const T*)src, (T*)dst);
  else
    k_step_geo<L, COLL, ST, 0, NPL, GF><<<grid, block, 0, s>>>(a, (const T*)src, (T*)dst);
}

// The group flags serve the 16-bit kernels, where a wave's bytes in flight are few and the
// dependent second round of bounce-back loads cost most (A/B on MI355X with and without the
// flags, two rounds each: sphere-array box 256^3 fp16 +7.5%; cylinder-array channel 4096^2
// bf16 +2.6%, fp16 +0.6%).  The fp32 kernels keep the mask-first order: holding the own
// slots of a flagged wave beside its pulled values in fp32 cost them 2.3% (D3Q19) and 0.7%
// (D2Q9).
template <class L, int COLL, int ST, int NPL>
int launch_geo_n(const KArgs& a0, int fmode, const void* src, void* dst, hipStream_t s) {
  const KArgs a = with_npl(a0, NPL);
  if constexpr (sizeof(typename Codec<ST>::T) == 2) {
    if (a.lgroups && a.links) {
      launch_geo_g<L, COLL, ST, NPL, true>(a, fmode, src, dst, s);
      return (int)hipGetLastError();
    }
  }
  launch_geo_g<L, COLL, ST, NPL, false>(a, fmode, src, dst, s);
  return (int)hipGetLastError();
}

template <class L, int COLL, int ST>
int launch_geo_t(const KArgs& a, int fmode, const void* src, void* dst, hipStream_t s) {
  constexpr int NPL = L::Q == 9 ? (sizeof(typename Codec<ST>::T) == 2 ? 4 : 1) : 1;
  if constexpr (NPL > 1) {
    if (a.nx >= kBlock * NPL) return launch_geo_n<L, COLL, ST, NPL>(a, fmode, src, dst, s);
  }
  return launch_geo_n<L, COLL, ST, 1>(a, fmode, src, dst, s);
}

template <class L, int COLL>
int launch_geo_s(const KArgs& a, int storage, int fmode, const void* src, void* dst,
                 hipStream_t s) {
  switch (storage) {
    case PYLBM_F32: return launch_geo_t<L, COLL, PYLBM_F32>(a, fmode, src, dst, s);
    case PYLBM_F16: return launch_geo_t<L, COLL, PYLBM_F16>(a, fmode, src, dst, s);
    case PYLBM_BF16: return launch_geo_t<L, COLL, PYLBM_BF16>(a, fmode, src, dst, s);
    case PYLBM_F32S: return launch_geo_t<L, COLL, PYLBM_F32S>(a, fmode, src, dst, s);
  }
  return PYLBM_EINVAL;
}

// Host-side dispatcher (two-buffer pattern only); instantiated per lattice in
// geo_d2q9.hip / geo_d3q19.hip.  The caller guarantees that the population array spans
// less than 4 GiB (single-descriptor addressing).
template <class L>
int launch_step_geo(const KArgs& a, int coll, int storage, int fmode, const void* src,
                    void* dst, hipStream_t s) {
  switch (coll) {
    case PYLBM_BGK: return launch_geo_s<L, PYLBM_BGK>(a, storage, fmode, src, dst, s);
    case PYLBM_TRT: return launch_geo_s<L, PYLBM_TRT>(a, storage, fmode, src, dst, s);
    case PYLBM_MRT:
      if constexpr (L::Q == 9)
        return launch_geo_s<L, PYLBM_MRT>(a, storage, fmode, src, dst, s);
      else
        return PYLBM_ENOTSUP;
  }
  return PYLBM_EINVAL;
}

// Group flags of a link mask: one wave per group of 64 consecutive nodes of a row; the word
// is 1 when a node of the group has a solid source or sits on a wall of a non-periodic x.
template <class L>
__global__ void __launch_bounds__(256)
k_link_groups(const LinkMask<L>* links, uint32_t* groups, int nx, int ngroups, int gpr,
              int perx) {
  const int g = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));  // uniform per wave
  if (g >= ngroups) return;
  const int row = g / gpr;
  const int x = (g - row * gpr) * 64 + (int)(threadIdx.x & 63);
  bool need = false;
  if (x < nx) {
    need = links[(size_t)row * nx + x] != 0;
    if (!perx && (x == 0 || x == nx - 1)) need = true;
  }
  const uint64_t any = __ballot(need);
  if ((threadIdx.x & 63) == 0) groups[g] = any != 0 ? 1u : 0u;
}

template <class L>
int launch_link_groups(const void* links, uint32_t* groups, int nx, int rows, int perx,
                       hipStream_t s) {
  const int gpr = (nx + 63) / 64;
  const int ngroups = rows * gpr;
  const dim3 grid((unsigned)((ngroups + 3) / 4)), block(256);
  k_link_groups<L><<<grid, block, 0, s>>>(reinterpret_cast<const LinkMask<L>*>(links), groups,
                                          nx, ngroups, gpr, perx);
  return (int)hipGetLastError();
}

template <class L>
int launch_links(const uint8_t* solid, void* links, int nx, int ny, int nz, hipStream_t s) {
  const int N = nx * ny * nz;
  const dim3 grid((unsigned)((N + 255) / 256)), block(256);
  k_links<L><<<grid, block, 0, s>>>(solid, reinterpret_cast<LinkMask<L>*>(links), nx, ny, nz,
                                    N);
  return (int)hipGetLastError();
}

}  // namespace plbm
