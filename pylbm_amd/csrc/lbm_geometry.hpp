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

#include "lbm_access.hpp"

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

template <class L, int COLL, int ST, int FM>
__global__ void __launch_bounds__(kBlock)
k_step_geo(const KArgs a, const typename Codec<ST>::T* A, typename Codec<ST>::T* B) {
  using C = Codec<ST>;
  using T = typename C::T;
  using M = LinkMask<L>;
  constexpr int Q = L::Q;
  constexpr uint32_t TS = sizeof(T);

  Lane l;
  if (!lane_setup<TS>(a, remap_block((int)blockIdx.x, a.nblk, a.remap), l)) return;
  const size_t plane = (size_t)a.N * TS;
  const char* Ab = reinterpret_cast<const char*>(A);
  char* Bb = reinterpret_cast<char*>(B);

  uint32_t bb = 0;
  if (a.links) bb = (uint32_t)ld_at<M>(reinterpret_cast<const char*>(a.links), l.n * sizeof(M));
  const float uin = a.inlet ? a.inlet[l.y] : 0.0f;  // uniform: one row per workgroup

  T t[Q];
  float add[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int cx = L::c(i, 0);
    add[i] = 0.0f;
    if (i == 0) {
      t[0] = ld_pl<T, true>(Ab, plane, 0, 0, l.o0);
      continue;
    }
    const RowLink r = row_link<L>(a, l.y, l.z, i, -1);
    const uint32_t offx = cx > 0 ? l.om : (cx < 0 ? l.op : l.o0);
    t[i] = ld_pl<T, true>(Ab, plane, r.wall ? L::opp(i) : i,
                          r.wall ? 0 : r.delta * (ptrdiff_t)TS, r.wall ? l.o0 : offx);
    if (cx != 0 && r.wall && r.top) add[i] = (float)(6.0 * L::w(i) * cx) * a.ulid;
    // inlet: crosses the x = 0 wall and no later (y / z) wall
    if (L::D == 2 && cx == 1 && !r.wall)
      add[i] = l.wall_lo ? (float)(6.0 * L::w(i)) * uin : 0.0f;
  }
  if (l.wall_lo) bb |= x_bits<L>(+1);
  if (l.wall_hi) bb |= x_bits<L>(-1);
  if (bb != 0) {
    T tw[Q];
#pragma unroll
    for (int i = 1; i < Q; ++i) tw[i] = ld_pl<T, true>(Ab, plane, L::opp(i), 0, l.o0);
#pragma unroll
    for (int i = 1; i < Q; ++i)
      if ((bb >> i) & 1u) t[i] = tw[i];
  }
  float h[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) h[i] = C::ld(t[i]) + (C::shifted ? 0.0f : -(float)L::w(i)) + add[i];

  if (L::D == 2 && a.outlet && l.x == a.nx - 1) {
    // rho = 1 equilibrium at the velocity of the stored column x - 1 (no force correction)
    T tc[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) tc[i] = ld_pl<T, true>(Ab, plane, i, 0, l.om);
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

  float F[3] = {0.f, 0.f, 0.f};
  if (FM == 1) {
    F[0] = a.fx;
    F[1] = a.fy;
    F[2] = a.fz;
  } else if (FM == 2) {
    const char* ffb = reinterpret_cast<const char*>(a.ff);
    const size_t fplane = (size_t)a.N * sizeof(float);
    const uint32_t of = l.n * 4u;
    F[0] = ld_at<float>(ffb, of);
    F[1] = ld_at<float>(ffb + fplane, of);
    F[2] = (L::D == 3) ? ld_at<float>(ffb + 2 * fplane, of) : 0.0f;
  }
  float d[Q];
  collide_node<L, COLL, FM != 0>(h, d, F, a.r0, a.r1, a.r2, a.r3);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const float v = C::shifted ? d[i] : d[i] + (float)L::w(i);
    st_pl<2, true>(Bb, plane, i, 0, l.o0, C::st(v));
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

template <class L, int COLL, int ST>
int launch_geo_t(const KArgs& a, int fmode, const void* src, void* dst, hipStream_t s) {
  using T = typename Codec<ST>::T;
  const dim3 grid((unsigned)a.nblk), block(kBlock);
  if (fmode == 2)
    k_step_geo<L, COLL, ST, 2><<<grid, block, 0, s>>>(a, (const T*)src, (T*)dst);
  else if (fmode == 1)
    k_step_geo<L, COLL, ST, 1><<<grid, block, 0, s>>>(a, (const T*)src, (T*)dst);
  else
    k_step_geo<L, COLL, ST, 0><<<grid, block, 0, s>>>(a, (const T*)src, (T*)dst);
  return (int)hipGetLastError();
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

template <class L>
int launch_links(const uint8_t* solid, void* links, int nx, int ny, int nz, hipStream_t s) {
  const int N = nx * ny * nz;
  const dim3 grid((unsigned)((N + 255) / 256)), block(256);
  k_links<L><<<grid, block, 0, s>>>(solid, reinterpret_cast<LinkMask<L>*>(links), nx, ny, nz,
                                    N);
  return (int)hipGetLastError();
}

}  // namespace plbm
