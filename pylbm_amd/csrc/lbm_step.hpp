// The fused stream-collide kernel for gfx950, templated over lattice, collision, storage,
// streaming pattern and force mode.
//
// Semantics (what is read and written per node) are those of the reference's fused kernels
// (pylbm/codegen.py:10-16 for the patterns, pylbm/backends/cuda/lbm.cu:71-98 and 165-198
// for pull / bounce-back / Ladd / AA push); the structure is this file's own:
//
//   * one lane per node, one workgroup = 256 lanes = one contiguous segment of one x-row
//     (fixed y, z), so every wave64 access to a plane covers 64 consecutive, naturally
//     aligned elements;
//   * the y/z part of every link (row wrap, wall, lid) is decided once per workgroup on the
//     scalar unit, and folded into one 64-bit base pointer per direction held in SGPRs;
//   * the lane part of every address is one of three 32-bit byte offsets (own node, x-1,
//     x+1, with the periodic wrap already applied) shared by all directions; loads and
//     stores are raw buffer instructions on a descriptor built from the uniform base, so
//     interior lanes spend no vector instructions on addressing;
//   * only the lanes at x = 0 / x = nx-1 of a row with x walls take a divergent branch that
//     reads (or, in the AA odd step, writes) their own bounce-back slot;
//   * a 1-D grid of row segments, remapped so each XCD owns a contiguous slab;
//   * stores of the two-buffer and AA-even patterns are non-temporal (written once, not
//     re-read within the step).
// One node per lane with natural alignment measured faster on MI355X than 2, 4 or 8 nodes
// per lane with unaligned 8/16-byte accesses along x (those kernels were removed).
#pragma once

#include "lbm_common.hpp"

namespace plbm {

enum ForceMode { kForceNone = 0, kForceUniform = 1, kForceField = 2 };

// Raw buffer access at base + off: base is uniform (a 4-SGPR descriptor, no bounds limit:
// every offset the kernels form is in range by construction), off a 32-bit lane offset.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc_at(const void* base) {
  // word 3 = 0x00020000: 32-bit data format, as gfx9-family raw buffers require
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)0xFFFFFFFF,
                                           0x00020000);
}

template <class T>
__device__ __forceinline__ T ld_at(const char* base, uint32_t off) {
  if constexpr (sizeof(T) == 4)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rsrc_at(base), off, 0, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b16(rsrc_at(base), off, 0, 0));
}

// AUX = 2 marks the store non-temporal (nt): written once, not re-read within the step.
template <int AUX, class T>
__device__ __forceinline__ void st_at(char* base, uint32_t off, T v) {
  if constexpr (sizeof(T) == 4)
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), rsrc_at(base), off, 0,
                                          AUX);
  else
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, v), rsrc_at(base), off, 0,
                                          AUX);
}

template <class L, int COLL, int ST, int PAT, int FM>
__global__ void __launch_bounds__(kBlock)
k_step(const KArgs a, const typename Codec<ST>::T* A, typename Codec<ST>::T* B) {
  using C = Codec<ST>;
  using T = typename C::T;
  constexpr int Q = L::Q;
  constexpr uint32_t TS = sizeof(T);

  const int b = remap_block((int)blockIdx.x, a.nblk, a.remap);
  const int row = (int)a.div_nseg.div((uint32_t)b);
  const int seg = b - row * a.nseg;
  const int z = (int)a.div_ny.div((uint32_t)row);
  const int y = row - z * a.ny;
  const int x = seg * kBlock + (int)threadIdx.x;
  if (x >= a.nx) return;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t N = (size_t)a.N;
  const bool perx = a.per & 1, pery = (a.per >> 1) & 1, perz = (a.per >> 2) & 1;

  // lane byte offsets: own node, source x-1 (pull along c_x = +1), source x+1 (c_x = -1)
  const uint32_t n = (uint32_t)(x + nx * (y + ny * z));
  int dxm1 = -1, dxp1 = 1;      // element offsets to x-1 and x+1 (wrapped)
  bool wall_lo = false, wall_hi = false;  // lane sits on an x wall (x = 0 / x = nx-1)
  if (x == 0) {
    if (perx) dxm1 = nx - 1;
    else { dxm1 = 0; wall_lo = true; }
  }
  if (x == nx - 1) {
    if (perx) dxp1 = -(nx - 1);
    else { dxp1 = 0; wall_hi = true; }
  }
  const uint32_t o0 = n * TS;
  const uint32_t om = (uint32_t)((int)n + dxm1) * TS;
  const uint32_t op = (uint32_t)((int)n + dxp1) * TS;
  const char* Ab = reinterpret_cast<const char*>(A);
  char* Bb = reinterpret_cast<char*>(B);
  const size_t plane = N * TS;

  // ---------------------------------------------------------------- gather (pull)
  // Every lane issues its Q loads unconditionally, back to back (links through a y/z wall
  // select their bounce-back slot with a uniform select, not a branch).  Lanes on an x wall
  // load a valid placeholder there and fetch their bounce-back values afterwards in one
  // divergent block, into separate registers, so no load waits on another.
  T t[Q];
  float lid[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int cx = L::c(i, 0), cy = L::c(i, 1), cz = L::c(i, 2);
    const int own_bb = (PAT == PYLBM_AB) ? L::opp(i) : i;  // bounce-back slot at own node
    const int far = (PAT == PYLBM_AB) ? i : L::opp(i);     // slot read at the neighbour
    lid[i] = 0.0f;
    if (PAT == PYLBM_AA_EVEN || i == 0) {
      const int slot = (PAT == PYLBM_AA_ODD) ? L::opp(i) : i;
      t[i] = ld_at<T>(Ab + slot * plane, o0);
      continue;
    }
    // row part of the link (uniform over the workgroup)
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
    const ptrdiff_t rdelta = (ptrdiff_t)nx * (sy - y) + (ptrdiff_t)nx * ny * (sz - z);
    const char* base = rwall ? Ab + own_bb * plane : Ab + far * plane + rdelta * (ptrdiff_t)TS;
    const uint32_t offx = cx > 0 ? om : (cx < 0 ? op : o0);
    t[i] = ld_at<T>(base, rwall ? o0 : offx);
    if (cx != 0) lid[i] = (rwall && top) ? (float)(6.0 * L::w(i) * cx) * a.ulid : 0.0f;
  }
  if (wall_lo || wall_hi) {
    T tw[Q];
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      if (PAT == PYLBM_AA_EVEN || L::c(i, 0) == 0) continue;
      tw[i] = ld_at<T>(Ab + ((PAT == PYLBM_AB) ? L::opp(i) : i) * plane, o0);
    }
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      if (PAT == PYLBM_AA_EVEN || L::c(i, 0) == 0) continue;
      // an x-wall link bounces back unless the same link already crossed a y/z wall
      // (then t[i] holds that bounce-back value already: identical slot)
      if ((L::c(i, 0) > 0 && wall_lo) || (L::c(i, 0) < 0 && wall_hi)) t[i] = tw[i];
    }
  }
  float h[Q];  // deviations f - w (shifted storages hold them; fp32 subtracts exactly)
#pragma unroll
  for (int i = 0; i < Q; ++i)
    h[i] = C::ld(t[i]) + (C::shifted ? 0.0f : -(float)L::w(i)) + lid[i];

  // ---------------------------------------------------------------- collide
  float F[3] = {0.f, 0.f, 0.f};
  if (FM == kForceUniform) {
    F[0] = a.fx;
    F[1] = a.fy;
    F[2] = a.fz;
  } else if (FM == kForceField) {
    const char* ffb = reinterpret_cast<const char*>(a.ff);
    const size_t fplane = N * sizeof(float);
    F[0] = ld_at<float>(ffb, n * 4u);
    F[1] = ld_at<float>(ffb + fplane, n * 4u);
    F[2] = (L::D == 3) ? ld_at<float>(ffb + 2 * fplane, n * 4u) : 0.0f;
  }
  float d[Q];
  collide_node<L, COLL, FM != kForceNone>(h, d, F, a.r0, a.r1, a.r2, a.r3);

  // ---------------------------------------------------------------- scatter
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const float v = C::shifted ? d[i] : d[i] + (float)L::w(i);
    if (PAT == PYLBM_AB) {
      st_at<2>(Bb + i * plane, o0, C::st(v));
      continue;
    }
    if (PAT == PYLBM_AA_EVEN) {
      st_at<2>(Bb + L::opp(i) * plane, o0, C::st(v));
      continue;
    }
    if (i == 0) {
      st_at<0>(Bb, o0, C::st(v));
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
    char* own = Bb + L::opp(i) * plane;
    if (rwall) {
      float vb = v;
      if (cx != 0 && top) vb = v + (float)(-6.0 * L::w(i) * cx) * a.ulid;
      st_at<0>(own, o0, C::st(vb));
      continue;
    }
    const ptrdiff_t rdelta = (ptrdiff_t)nx * (ty - y) + (ptrdiff_t)nx * ny * (tz - z);
    char* base = Bb + i * plane + rdelta * (ptrdiff_t)TS;
    const uint32_t off = cx > 0 ? op : (cx < 0 ? om : o0);  // push target x + c_x
    if ((cx > 0 && wall_hi) || (cx < 0 && wall_lo)) st_at<0>(own, o0, C::st(v));
    else st_at<0>(base, off, C::st(v));
  }
}

template <class L, int COLL, int ST, int PAT>
int launch_step_t(const KArgs& a, int fmode, const void* src, void* dst, hipStream_t s) {
  using T = typename Codec<ST>::T;
  const dim3 grid((unsigned)a.nblk), block(kBlock);
  if (fmode == kForceField)
    k_step<L, COLL, ST, PAT, kForceField><<<grid, block, 0, s>>>(a, (const T*)src, (T*)dst);
  else if (fmode == kForceUniform)
    k_step<L, COLL, ST, PAT, kForceUniform><<<grid, block, 0, s>>>(a, (const T*)src, (T*)dst);
  else
    k_step<L, COLL, ST, PAT, kForceNone><<<grid, block, 0, s>>>(a, (const T*)src, (T*)dst);
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
