// Memory access of the fused kernels (step and adjoint): raw buffer loads / stores on
// uniform bases, the row-segment decomposition of a workgroup, per-lane x offsets, the
// y/z part of a link, and the pull gather.
//
//   * one lane per node, one workgroup = 256 lanes = one contiguous segment of one x-row
//     (fixed y, z), so every wave64 access to a plane covers 64 consecutive, naturally
//     aligned elements;
//   * the y/z part of every link (row wrap, wall, lid) is decided once per workgroup on the
//     scalar unit (the block index is split with an exact reciprocal multiply, FastDiv,
//     since a run-time integer division would run on the vector unit) and folded into one
//     base address per direction;
//   * the lane part of every address is one of three 32-bit byte offsets (own node, x-1,
//     x+1, periodic wrap applied) shared by all directions, and accesses are raw buffer
//     instructions on a descriptor built from the uniform base, so interior lanes spend no
//     vector instructions on addressing;
//   * the gather issues all its loads back to back; lanes on an x wall read their
//     bounce-back values afterwards, in one divergent block, into separate registers.
#pragma once

#include "lbm_common.hpp"

namespace plbm {

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

// AUX = 2 marks the store non-temporal (nt): written once, not re-read by the same kernel.
template <int AUX, class T>
__device__ __forceinline__ void st_at(char* base, uint32_t off, T v) {
  if constexpr (sizeof(T) == 4)
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), rsrc_at(base), off, 0,
                                          AUX);
  else
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, v), rsrc_at(base), off, 0,
                                          AUX);
}

// Moment record of one node: (rho, u_x, u_y, u_z) fp32, 16 bytes at byte offset off of
// the node-major record array (one 16-byte access per lane, so a wave moves 1 KiB).
typedef unsigned int rec4_t __attribute__((ext_vector_type(4)));

template <int AUX>
__device__ __forceinline__ void st_rec(float* rec, uint32_t off, const float (&mo)[4]) {
  rec4_t v;
  v.x = __builtin_bit_cast(uint32_t, mo[0]);
  v.y = __builtin_bit_cast(uint32_t, mo[1]);
  v.z = __builtin_bit_cast(uint32_t, mo[2]);
  v.w = __builtin_bit_cast(uint32_t, mo[3]);
  __builtin_amdgcn_raw_buffer_store_b128(v, rsrc_at(rec), off, 0, AUX);
}

// The record of node n, read once (non-temporal).  A global load: this compiler lowers
// __builtin_amdgcn_raw_buffer_load_b128 to a single-dword load and leaves the other three
// components undefined (the b128 store above is lowered correctly, to buffer_store_dwordx4).
typedef float rec4f_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void ld_rec_nt(const float* rec, uint32_t n, float (&mo)[4]) {
  const rec4f_t v = __builtin_nontemporal_load(reinterpret_cast<const rec4f_t*>(rec) + n);
  mo[0] = v.x;
  mo[1] = v.y;
  mo[2] = v.z;
  mo[3] = v.w;
}

// Access to element (plane slot, node) of a state array: base + slot * plane + delta + off
// with delta a uniform byte offset (the row shift of a link, |delta| < plane) and off the
// lane's byte offset.  ADDR chooses how the uniform part reaches the hardware:
//   1  the whole array spans less than 4 GiB: one descriptor for the array, and
//      slot * plane + delta (non-negative: |delta| < plane for every slot >= 1, slot 0
//      never shifts) in the scalar soffset operand;
//   G  (>= 2) larger arrays: the planes form groups of G slots, group g read through one
//      descriptor based one plane before its first slot (at plane 0 for g = 0), so that
//      soffset = (slot - base slot) * plane + delta stays non-negative and soffset + off
//      stays below (G + 2) planes, which the host keeps under 4 GiB; a kernel then holds
//      ceil(q / G) descriptors per array instead of one per plane and row;
//   0  one descriptor per access at its plane-and-row base (any size).
template <int ADDR>
__device__ __forceinline__ void plane_ref(const char* A, size_t plane, int slot, ptrdiff_t delta,
                                          const char*& base, uint32_t& so) {
  if constexpr (ADDR == 1) {
    base = A;
    so = (uint32_t)(slot * plane + delta);
  } else {
    const int g = slot / ADDR;
    const int first = g == 0 ? 0 : g * ADDR - 1;
    base = A + (size_t)first * plane;
    so = (uint32_t)((slot - first) * plane + delta);
  }
}

// LA: cache policy bits of the load (0 default; 2 non-temporal)
template <class T, int ADDR, int LA = 0>
__device__ __forceinline__ T ld_pl(const char* A, size_t plane, int slot, ptrdiff_t delta,
                                   uint32_t off) {
  if constexpr (ADDR >= 1) {
    const char* base;
    uint32_t so;
    plane_ref<ADDR>(A, plane, slot, delta, base, so);
    if constexpr (sizeof(T) == 4)
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rsrc_at(base), off, so, LA));
    else
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b16(rsrc_at(base), off, so, LA));
  } else {
    return ld_at<T>(A + slot * plane + delta, off);
  }
}

template <int AUX, int ADDR, class T>
__device__ __forceinline__ void st_pl(char* A, size_t plane, int slot, ptrdiff_t delta,
                                      uint32_t off, T v) {
  if constexpr (ADDR >= 1) {
    const char* base;
    uint32_t so;
    plane_ref<ADDR>(A, plane, slot, delta, base, so);
    char* b = const_cast<char*>(base);
    if constexpr (sizeof(T) == 4)
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), rsrc_at(b), off, so,
                                            AUX);
    else
      __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, v), rsrc_at(b), off, so,
                                            AUX);
  } else {
    st_at<AUX>(A + slot * plane + delta, off, v);
  }
}

// A lane's node and the element indices of its own node (n) and its x-1 / x+1 neighbours
// in the same row (nm, np), with the periodic wrap applied; on an x wall the neighbour
// index is the own node (a valid placeholder) and wall_lo / wall_hi is set.  o0 / om / op
// are the byte offsets of the same elements for element size TS.
struct Lane {
  int x, y, z;
  uint32_t n, nm, np, o0, om, op;
  bool wall_lo, wall_hi;
};

// Row (y, z) and segment index of the calling wave in workgroup blk, uniform per wave (the
// scalar unit computes it).  A workgroup is one segment of kBlock lanes of a row, or, for
// rows of at most 128 nodes, 2 or 4 whole rows of 128 or 64 lanes each, so that short rows
// leave no wave idle.  False for the waves of the last workgroup past the final row.
__device__ __forceinline__ bool block_row(const KArgs& a, int blk, int& y, int& z, int& seg) {
  int row;
  if (a.rpw_log2 == 0) {
    row = (int)a.div_nseg.div((uint32_t)blk);
    seg = blk - row * a.nseg;
  } else {
    row = (blk << a.rpw_log2) +
          __builtin_amdgcn_readfirstlane((int)threadIdx.x >> a.sw_log2);
    seg = 0;
  }
  z = (int)a.div_ny.div((uint32_t)row);
  y = row - z * a.ny;
  return row < a.rows;
}

// Lane index within its row segment.
__device__ __forceinline__ int seg_lane(const KArgs& a) {
  return (int)threadIdx.x & ((1 << a.sw_log2) - 1);
}

// Lane of node (x, y, z); false when x lies past the end of the row.
template <uint32_t TS>
__device__ __forceinline__ bool lane_at(const KArgs& a, int y, int z, int x, Lane& l) {
  l.x = x;
  l.y = y;
  l.z = z;
  const int nx = a.nx;
  const bool perx = a.per & 1;
  l.n = (uint32_t)(l.x + nx * (l.y + a.ny * l.z));
  int dxm1 = -1, dxp1 = 1;
  l.wall_lo = l.wall_hi = false;
  if (l.x == 0) {
    if (perx) dxm1 = nx - 1;
    else { dxm1 = 0; l.wall_lo = true; }
  }
  if (l.x == nx - 1) {
    if (perx) dxp1 = -(nx - 1);
    else { dxp1 = 0; l.wall_hi = true; }
  }
  l.nm = (uint32_t)((int)l.n + dxm1);
  l.np = (uint32_t)((int)l.n + dxp1);
  l.o0 = l.n * TS;
  l.om = l.nm * TS;
  l.op = l.np * TS;
  return l.x < nx;
}

// One node per lane: the lane's node in workgroup blk (false: no node).
template <uint32_t TS>
__device__ __forceinline__ bool lane_setup(const KArgs& a, int blk, Lane& l) {
  int y, z, seg;
  if (!block_row(a, blk, y, z, seg)) return false;
  return lane_at<TS>(a, y, z, seg * kBlock + seg_lane(a), l);
}

// The y/z part of link i seen from row (y, z) towards y + sign c_y, z + sign c_z
// (sign = -1: the pull source, +1: the push target): wall (the link leaves the box through
// a y or z wall), top (through the lid: the high wall of the last axis) and the element
// offset of the wrapped target row relative to the own row.  Uniform over a workgroup.
struct RowLink {
  bool wall, top;
  ptrdiff_t delta;
};

template <class L>
__device__ __forceinline__ RowLink row_link(const KArgs& a, int y, int z, int i, int sign) {
  const int cy = sign * L::c(i, 1), cz = sign * L::c(i, 2);
  const int ny = a.ny, nz = a.nz;
  const bool pery = (a.per >> 1) & 1, perz = (a.per >> 2) & 1;
  int ty = y + cy, tz = z + cz;
  RowLink r{false, false, 0};
  if (cy != 0 && (ty < 0 || ty >= ny)) {
    if (pery) ty += (ty < 0) ? ny : -ny;
    else { r.wall = true; if (L::D == 2 && ty >= ny) r.top = true; }
  }
  if (L::D == 3 && cz != 0 && (tz < 0 || tz >= nz)) {
    if (perz) tz += (tz < 0) ? nz : -nz;
    else { r.wall = true; if (tz >= nz) r.top = true; }
  }
  r.delta = (ptrdiff_t)a.nx * (ty - y) + (ptrdiff_t)a.nx * ny * (tz - z);
  return r;
}

// Pull gather for pattern PAT (pylbm/codegen.py:10-16): h_i = f_i - w_i of the population
// that streams into the lane's node, with half-way bounce-back and Ladd's lid term.
//   AB      : read slot i at x - c_i, or own slot opp(i) on a wall link
//   AA odd  : read slot opp(i) at x - c_i, or own slot i on a wall link
//   AA even : read own slot i
// gather = gather_issue (every load, back to back) + gather_finish (x-wall bounce-back in
// one divergent block, decode); split so that a lane serving several nodes issues the
// loads of all of them before it waits for any.
// LA: cache policy of the loads of links with no x component, LAX: of the x-shifted links
// (a line an x-shifted link reads is also read by the neighbouring wave or row segment,
// which may find it in the L2; the lines of a link with no x component belong to this
// segment alone)
template <class L, int ST, int PAT, int ADDR, int LA = 0, int LAX = LA>
__device__ __forceinline__ void gather_issue(const KArgs& a, const Lane& l, const char* Ab,
                                             size_t plane, typename Codec<ST>::T (&t)[L::Q],
                                             float (&lid)[L::Q]) {
  using T = typename Codec<ST>::T;
  constexpr int Q = L::Q;
  constexpr uint32_t TS = sizeof(T);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int cx = L::c(i, 0);
    const int own_bb = (PAT == PYLBM_AB) ? L::opp(i) : i;
    const int far = (PAT == PYLBM_AB) ? i : L::opp(i);
    lid[i] = 0.0f;
    if (PAT == PYLBM_AA_EVEN || i == 0) {
      const int slot = (PAT == PYLBM_AA_ODD) ? L::opp(i) : i;
      t[i] = ld_pl<T, ADDR, LA>(Ab, plane, slot, 0, l.o0);
      continue;
    }
    const RowLink r = row_link<L>(a, l.y, l.z, i, -1);
    const uint32_t offx = cx > 0 ? l.om : (cx < 0 ? l.op : l.o0);
    if (cx == 0)
      t[i] = ld_pl<T, ADDR, LA>(Ab, plane, r.wall ? own_bb : far,
                             r.wall ? 0 : r.delta * (ptrdiff_t)TS, r.wall ? l.o0 : offx);
    else
      t[i] = ld_pl<T, ADDR, LAX>(Ab, plane, r.wall ? own_bb : far,
                              r.wall ? 0 : r.delta * (ptrdiff_t)TS, r.wall ? l.o0 : offx);
    if (cx != 0) lid[i] = (r.wall && r.top) ? (float)(6.0 * L::w(i) * cx) * a.ulid : 0.0f;
  }
}

template <class L, int ST, int PAT, int ADDR>
__device__ __forceinline__ void gather_finish(const Lane& l, const char* Ab, size_t plane,
                                              typename Codec<ST>::T (&t)[L::Q],
                                              const float (&lid)[L::Q], float (&h)[L::Q]) {
  using C = Codec<ST>;
  using T = typename C::T;
  constexpr int Q = L::Q;
  if (PAT != PYLBM_AA_EVEN && (l.wall_lo || l.wall_hi)) {
    T tw[Q];
#pragma unroll
    for (int i = 1; i < Q; ++i)
      if (L::c(i, 0) != 0)
        tw[i] = ld_pl<T, ADDR>(Ab, plane, (PAT == PYLBM_AB) ? L::opp(i) : i, 0, l.o0);
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      // an x-wall link bounces back; if it also crossed a y/z wall, t[i] already holds the
      // same own-slot value
      if ((L::c(i, 0) > 0 && l.wall_lo) || (L::c(i, 0) < 0 && l.wall_hi)) t[i] = tw[i];
    }
  }
#pragma unroll
  for (int i = 0; i < Q; ++i) h[i] = C::ld(t[i]) + (C::shifted ? 0.0f : -(float)L::w(i)) + lid[i];
}

template <class L, int ST, int PAT, int ADDR>
__device__ __forceinline__ void gather(const KArgs& a, const Lane& l, const char* Ab,
                                       size_t plane, float (&h)[L::Q]) {
  typename Codec<ST>::T t[L::Q];
  float lid[L::Q];
  gather_issue<L, ST, PAT, ADDR>(a, l, Ab, plane, t, lid);
  gather_finish<L, ST, PAT, ADDR>(l, Ab, plane, t, lid, h);
}

}  // namespace plbm
