// The fused stream-collide kernel for gfx950, templated over lattice, collision, storage,
// streaming pattern and force mode.
//
// Semantics (what is read and written per node) are those of the reference's fused kernels
// (pylbm/codegen.py:10-16 for the patterns, pylbm/backends/cuda/lbm.cu:71-98 and 165-198
// for pull / bounce-back / Ladd / AA push); the structure is this file's own:
//
//   * one workgroup = 256 lanes = one contiguous segment of one x-row (fixed y, z), so the
//     y/z wall and wrap decisions of a direction are uniform over the workgroup and the
//     lanes of a wave touch consecutive addresses of every plane (coalesced wave64 access);
//   * a 1-D grid of row segments, optionally remapped so each XCD owns a contiguous slab;
//   * each plane base is a 64-bit pointer, in-plane indices are 32-bit;
//   * stores of the two-buffer and AA-even patterns are non-temporal (written once, not
//     re-read within the step).
#pragma once

#include "lbm_common.hpp"

namespace plbm {

enum ForceMode { kForceUniform = 0, kForceField = 1 };

template <class L, int COLL, int ST, int PAT, int FM>
__global__ void __launch_bounds__(kBlock)
k_step(const KArgs a, const typename Codec<ST>::T* A, typename Codec<ST>::T* B) {
  using C = Codec<ST>;
  using T = typename C::T;
  constexpr int Q = L::Q;

  const int b = remap_block((int)blockIdx.x, a.nblk, a.remap);
  const int seg = b % a.nseg;
  const int row = b / a.nseg;
  const int y = row % a.ny;
  const int z = row / a.ny;
  const int x = seg * kBlock + (int)threadIdx.x;
  if (x >= a.nx) return;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int n = x + nx * (y + ny * z);
  const size_t N = (size_t)a.N;
  const bool perx = a.per & 1, pery = (a.per >> 1) & 1, perz = (a.per >> 2) & 1;

  // ---------------------------------------------------------------- gather (pull)
  float f[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int cx = L::c(i, 0), cy = L::c(i, 1), cz = L::c(i, 2);
    // f holds deviations h = f - w: shifted storages keep them, fp32 subtracts exactly
    const float shift = C::shifted ? 0.0f : -(float)L::w(i);
    if (PAT == PYLBM_AA_EVEN || i == 0) {
      const int slot = (PAT == PYLBM_AA_ODD) ? L::opp(i) : i;
      f[i] = C::ld(A[(size_t)slot * N + n]) + shift;
      continue;
    }
    int sx = x - cx, sy = y - cy, sz = z - cz;
    bool wall = false, top = false;
    if (cx != 0 && (sx < 0 || sx >= nx)) {
      if (perx) sx += (sx < 0) ? nx : -nx;
      else wall = true;
    }
    if (cy != 0 && (sy < 0 || sy >= ny)) {
      if (pery) sy += (sy < 0) ? ny : -ny;
      else { wall = true; if (L::D == 2 && sy >= ny) top = true; }
    }
    if (L::D == 3 && cz != 0 && (sz < 0 || sz >= nz)) {
      if (perz) sz += (sz < 0) ? nz : -nz;
      else { wall = true; if (sz >= nz) top = true; }
    }
    size_t idx;
    if (PAT == PYLBM_AB)
      idx = wall ? (size_t)L::opp(i) * N + n : (size_t)i * N + (sx + nx * (sy + ny * sz));
    else  // AA odd: the neighbour holds population i in slot opp(i); own slot i on a wall
      idx = wall ? (size_t)i * N + n : (size_t)L::opp(i) * N + (sx + nx * (sy + ny * sz));
    float v = C::ld(A[idx]) + shift;
    if (cx != 0 && top) v += (float)(6.0 * L::w(i) * cx) * a.ulid;
    f[i] = v;
  }

  // ---------------------------------------------------------------- collide
  float F[3] = {a.fx, a.fy, a.fz};
  if (FM == kForceField) {
    F[0] = a.ff[n];
    F[1] = a.ff[N + n];
    F[2] = (L::D == 3) ? a.ff[2 * N + n] : 0.0f;
  }
  const bool force = (FM == kForceField) || (F[0] != 0.f || F[1] != 0.f || F[2] != 0.f);
  float p[Q];
  collide_node<L, COLL>(f, p, F, force, a.r0, a.r1, a.r2, a.r3);

  // ---------------------------------------------------------------- scatter
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const float v = C::shifted ? p[i] : p[i] + (float)L::w(i);
    if (PAT == PYLBM_AB) {
      __builtin_nontemporal_store(C::st(v), &B[(size_t)i * N + n]);
    } else if (PAT == PYLBM_AA_EVEN) {
      __builtin_nontemporal_store(C::st(v), &B[(size_t)L::opp(i) * N + n]);
    } else if (i == 0) {
      B[n] = C::st(v);
    } else {  // AA odd: push to x + c_i, or bounce back into the own slot opp(i)
      const int cx = L::c(i, 0), cy = L::c(i, 1), cz = L::c(i, 2);
      int tx = x + cx, ty = y + cy, tz = z + cz;
      bool wall = false, top = false;
      if (cx != 0 && (tx < 0 || tx >= nx)) {
        if (perx) tx += (tx < 0) ? nx : -nx;
        else wall = true;
      }
      if (cy != 0 && (ty < 0 || ty >= ny)) {
        if (pery) ty += (ty < 0) ? ny : -ny;
        else { wall = true; if (L::D == 2 && ty >= ny) top = true; }
      }
      if (L::D == 3 && cz != 0 && (tz < 0 || tz >= nz)) {
        if (perz) tz += (tz < 0) ? nz : -nz;
        else { wall = true; if (tz >= nz) top = true; }
      }
      if (wall) {
        float vb = v;
        if (cx != 0 && top) vb = v + (float)(-6.0 * L::w(i) * cx) * a.ulid;
        B[(size_t)L::opp(i) * N + n] = C::st(vb);
      } else {
        B[(size_t)i * N + (tx + nx * (ty + ny * tz))] = C::st(v);
      }
    }
  }
}

// ---------------------------------------------------------------------------------------
// Vectorised variant: each lane owns VEC consecutive nodes along x and moves 16 bytes per
// plane per access (VEC = 4 fp32 or 8 16-bit values), so a wave64 access covers 1 KiB of a
// plane.  The x-shifted planes (c_x = +-1) are read / pushed with one 16-byte access at a
// one-element offset; the hardware serves the unaligned access from the cache lines the
// neighbouring lanes touch anyway.  Links that cross a y or z boundary are decided once per
// workgroup (every lane of a workgroup shares y and z).  Only the lanes holding x = 0 or
// x = nx-1 resolve x wrap / walls element by element.  Requires nx % VEC == 0.

template <class T, int V>
__device__ __forceinline__ void vld(float (&out)[V], const T* p, float shift) {
  T t[V];
  __builtin_memcpy(t, p, sizeof(t));
#pragma unroll
  for (int v = 0; v < V; ++v) out[v] = Codec_ld(t[v]) + shift;
}

template <class L, int COLL, int ST, int PAT, int FM, int VEC>
__global__ void __launch_bounds__(kBlock)
k_step_vec(const KArgs a, const typename Codec<ST>::T* A, typename Codec<ST>::T* B) {
  using C = Codec<ST>;
  using T = typename C::T;
  constexpr int Q = L::Q;

  const int b = remap_block((int)blockIdx.x, a.nblk, a.remap);
  const int seg = b % a.nseg;
  const int row = b / a.nseg;
  const int y = row % a.ny;
  const int z = row / a.ny;
  const int x0 = (seg * kBlock + (int)threadIdx.x) * VEC;
  if (x0 >= a.nx) return;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int n0 = x0 + nx * (y + ny * z);
  const size_t N = (size_t)a.N;
  const bool perx = a.per & 1, pery = (a.per >> 1) & 1, perz = (a.per >> 2) & 1;
  const bool edge = (x0 == 0) || (x0 + VEC == nx);

  float f[Q][VEC];  // deviations h = f - w (see the scalar kernel)
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int cx = L::c(i, 0), cy = L::c(i, 1), cz = L::c(i, 2);
    const float shift = C::shifted ? 0.0f : -(float)L::w(i);
    if (PAT == PYLBM_AA_EVEN || i == 0) {
      const int slot = (PAT == PYLBM_AA_ODD) ? L::opp(i) : i;
      vld<T, VEC>(f[i], A + (size_t)slot * N + n0, shift);
      continue;
    }
    // y / z part of the link: uniform over the workgroup
    int sy = y - cy, sz = z - cz;
    bool rwall = false, top = false;
    if (cy != 0 && (sy < 0 || sy >= ny)) {
      if (pery) sy += (sy < 0) ? ny : -ny;
      else { rwall = true; if (L::D == 2 && sy >= ny) top = true; }
    }
    if (L::D == 3 && cz != 0 && (sz < 0 || sz >= nz)) {
      if (perz) sz += (sz < 0) ? nz : -nz;
      else { rwall = true; if (sz >= nz) top = true; }
    }
    const int own_bb = (PAT == PYLBM_AB) ? L::opp(i) : i;   // bounce-back slot at own node
    const int far = (PAT == PYLBM_AB) ? i : L::opp(i);      // slot read at the neighbour
    if (rwall) {
      vld<T, VEC>(f[i], A + (size_t)own_bb * N + n0, shift);
      if (cx != 0 && top) {
        const float lid = (float)(6.0 * L::w(i) * cx) * a.ulid;
#pragma unroll
        for (int v = 0; v < VEC; ++v) f[i][v] += lid;
      }
      continue;
    }
    const T* rowp = A + (size_t)far * N + (n0 + nx * (sy - y) + nx * ny * (sz - z));
    if (cx == 0 || !edge) {
      vld<T, VEC>(f[i], rowp - cx, shift);
    } else {
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        const int sx = x0 + v - cx;
        T t;
        if (sx >= 0 && sx < nx) t = rowp[v - cx];
        else if (perx) t = rowp[v - cx + (sx < 0 ? nx : -nx)];
        else t = A[(size_t)own_bb * N + n0 + v];
        f[i][v] = Codec_ld(t) + shift;
      }
    }
  }

  float F[VEC][3];
#pragma unroll
  for (int v = 0; v < VEC; ++v) {
    F[v][0] = a.fx; F[v][1] = a.fy; F[v][2] = a.fz;
  }
  if (FM == kForceField) {
    float t[VEC];
    __builtin_memcpy(t, a.ff + n0, sizeof(t));
#pragma unroll
    for (int v = 0; v < VEC; ++v) F[v][0] = t[v];
    __builtin_memcpy(t, a.ff + N + n0, sizeof(t));
#pragma unroll
    for (int v = 0; v < VEC; ++v) F[v][1] = t[v];
    if (L::D == 3) {
      __builtin_memcpy(t, a.ff + 2 * N + n0, sizeof(t));
#pragma unroll
      for (int v = 0; v < VEC; ++v) F[v][2] = t[v];
    }
  }
  const bool uforce = (a.fx != 0.f || a.fy != 0.f || a.fz != 0.f);
#pragma unroll
  for (int v = 0; v < VEC; ++v) {
    float fn[Q], pn[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) fn[i] = f[i][v];
    collide_node<L, COLL>(fn, pn, F[v], FM == kForceField || uforce, a.r0, a.r1, a.r2, a.r3);
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i][v] = C::shifted ? pn[i] : pn[i] + (float)L::w(i);
  }

#pragma unroll
  for (int i = 0; i < Q; ++i) {
    T t[VEC];
#pragma unroll
    for (int v = 0; v < VEC; ++v) t[v] = C::st(f[i][v]);
    if (PAT == PYLBM_AB || PAT == PYLBM_AA_EVEN || i == 0) {
      const int slot = (PAT == PYLBM_AA_EVEN) ? L::opp(i) : i;
      T* dst = B + (size_t)slot * N + n0;
      if (PAT == PYLBM_AA_ODD) {
        __builtin_memcpy(dst, t, sizeof(t));
      } else {
        typedef T TV __attribute__((ext_vector_type(VEC)));
        TV tv;
#pragma unroll
        for (int v = 0; v < VEC; ++v) tv[v] = t[v];
        __builtin_nontemporal_store(tv, (TV*)dst);
      }
      continue;
    }
    // AA odd: push to x + c_i, or bounce back into the own slot opp(i)
    const int cx = L::c(i, 0), cy = L::c(i, 1), cz = L::c(i, 2);
    int ty = y + cy, tz = z + cz;
    bool rwall = false, top = false;
    if (cy != 0 && (ty < 0 || ty >= ny)) {
      if (pery) ty += (ty < 0) ? ny : -ny;
      else { rwall = true; if (L::D == 2 && ty >= ny) top = true; }
    }
    if (L::D == 3 && cz != 0 && (tz < 0 || tz >= nz)) {
      if (perz) tz += (tz < 0) ? nz : -nz;
      else { rwall = true; if (tz >= nz) top = true; }
    }
    T* own = B + (size_t)L::opp(i) * N + n0;
    if (rwall) {
      if (cx != 0 && top) {
        const float lid = (float)(-6.0 * L::w(i) * cx) * a.ulid;
#pragma unroll
        for (int v = 0; v < VEC; ++v) t[v] = C::st(f[i][v] + lid);
      }
      __builtin_memcpy(own, t, sizeof(t));
      continue;
    }
    T* rowp = B + (size_t)i * N + (n0 + nx * (ty - y) + nx * ny * (tz - z));
    if (cx == 0 || !edge) {
      __builtin_memcpy(rowp + cx, t, sizeof(t));
    } else {
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        const int tx = x0 + v + cx;
        if (tx >= 0 && tx < nx) rowp[v + cx] = t[v];
        else if (perx) rowp[v + cx + (tx < 0 ? nx : -nx)] = t[v];
        else own[v] = t[v];
      }
    }
  }
}

// Nodes per lane of the vectorised kernels that are instantiated for each storage type.
template <int ST> constexpr int vec_lo() { return sizeof(typename Codec<ST>::T) == 4 ? 2 : 4; }
template <int ST> constexpr int vec_hi() { return sizeof(typename Codec<ST>::T) == 4 ? 4 : 8; }

template <class L, int COLL, int ST, int PAT, int V>
int launch_vec(const KArgs& a0, bool field, const void* src, void* dst, hipStream_t s) {
  using T = typename Codec<ST>::T;
  KArgs a = a0;  // grid of the vectorised kernel: kBlock * V nodes per row segment
  a.nseg = (a.nx / V + kBlock - 1) / kBlock;
  a.nblk = a.nseg * a.ny * a.nz;
  a.remap = (a.nblk % 8 == 0) ? 1 : 0;
  const dim3 grid((unsigned)a.nblk), block(kBlock);
  if (field)
    k_step_vec<L, COLL, ST, PAT, kForceField, V><<<grid, block, 0, s>>>(a, (const T*)src,
                                                                       (T*)dst);
  else
    k_step_vec<L, COLL, ST, PAT, kForceUniform, V><<<grid, block, 0, s>>>(a, (const T*)src,
                                                                         (T*)dst);
  return (int)hipGetLastError();
}

template <class L, int COLL, int ST, int PAT>
int launch_step_t(const KArgs& a, bool field, const void* src, void* dst, hipStream_t s) {
  using T = typename Codec<ST>::T;
  if (a.vec == vec_hi<ST>()) return launch_vec<L, COLL, ST, PAT, vec_hi<ST>()>(a, field, src, dst, s);
  if (a.vec == vec_lo<ST>()) return launch_vec<L, COLL, ST, PAT, vec_lo<ST>()>(a, field, src, dst, s);
  const dim3 grid((unsigned)a.nblk), block(kBlock);
  if (field)
    k_step<L, COLL, ST, PAT, kForceField><<<grid, block, 0, s>>>(a, (const T*)src, (T*)dst);
  else
    k_step<L, COLL, ST, PAT, kForceUniform><<<grid, block, 0, s>>>(a, (const T*)src, (T*)dst);
  return (int)hipGetLastError();
}

template <class L, int COLL, int ST>
int launch_step_p(const KArgs& a, int pattern, bool field, const void* src, void* dst,
                  hipStream_t s) {
  switch (pattern) {
    case PYLBM_AB: return launch_step_t<L, COLL, ST, PYLBM_AB>(a, field, src, dst, s);
    case PYLBM_AA_ODD: return launch_step_t<L, COLL, ST, PYLBM_AA_ODD>(a, field, src, dst, s);
    case PYLBM_AA_EVEN: return launch_step_t<L, COLL, ST, PYLBM_AA_EVEN>(a, field, src, dst, s);
  }
  return PYLBM_EINVAL;
}

template <class L, int COLL>
int launch_step_s(const KArgs& a, int storage, int pattern, bool field, const void* src,
                  void* dst, hipStream_t s) {
  switch (storage) {
    case PYLBM_F32: return launch_step_p<L, COLL, PYLBM_F32>(a, pattern, field, src, dst, s);
    case PYLBM_F16: return launch_step_p<L, COLL, PYLBM_F16>(a, pattern, field, src, dst, s);
    case PYLBM_BF16: return launch_step_p<L, COLL, PYLBM_BF16>(a, pattern, field, src, dst, s);
    case PYLBM_F32S: return launch_step_p<L, COLL, PYLBM_F32S>(a, pattern, field, src, dst, s);
  }
  return PYLBM_EINVAL;
}

// Host-side dispatcher for one lattice; each lattice is instantiated in its own
// translation unit (step_d2q9.hip, step_d3q19.hip) so they compile in parallel.
template <class L>
int launch_step(const KArgs& a, int coll, int storage, int pattern, bool field,
                const void* src, void* dst, hipStream_t s) {
  switch (coll) {
    case PYLBM_BGK: return launch_step_s<L, PYLBM_BGK>(a, storage, pattern, field, src, dst, s);
    case PYLBM_TRT: return launch_step_s<L, PYLBM_TRT>(a, storage, pattern, field, src, dst, s);
    case PYLBM_MRT:
      if constexpr (L::Q == 9)
        return launch_step_s<L, PYLBM_MRT>(a, storage, pattern, field, src, dst, s);
      else
        return PYLBM_ENOTSUP;
  }
  return PYLBM_EINVAL;
}

}  // namespace plbm
