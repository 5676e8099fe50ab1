// Two time steps per launch (temporal blocking) of the two-buffer pull, on D2Q9 boxes
// periodic on both axes: src -> dst over two steps, the intermediate state kept on chip.
//
// A single step reads and writes the whole state (2 q s bytes per node); two chained steps
// do it twice.  Here step A's result never reaches memory:
//   * a workgroup (one lane per column) holds a tile of kPairCols columns and sweeps the
//     rows of one y-chunk, y0 - 1 .. y1: step A pulls row r from src (the ordinary gather)
//     and keeps its post-collision populations in LDS rings, encoded exactly as a stored
//     state; step B runs on row r - 1 once step A has done row r, pulling from the rings,
//     and stores to dst;
//   * step B of row q pulls direction i from row q - c_y,i, column x - c_x,i, so the rings
//     keep per column only what a later row still needs: the 3 components with c_y = -1 of
//     the newest A row, the 3 with c_y = 0 of the two newest, the 3 with c_y = +1 of the
//     three newest (18 values), for the tile's columns and one halo column on each side,
//     which lanes 0 and 1 of the first wave compute;
//   * step A also runs on the halo rows y0 - 1 and y1 (1 + 2 / Ly of the reads);
//   * step A's loads of row r are issued before step B of row r - 2 runs, so that they fly
//     while it computes; the y-chunks are short (pair_chunk_rows), and each XCD takes a
//     contiguous range of tiles.
// Per node and two steps: src read once (plus the halo), dst written once, half the bytes of
// two single steps.  Every value is the single steps' value: the same gather, decode and
// collision on the same encoded populations.
#pragma once

#include <atomic>

#include "lbm_step.hpp"

namespace plbm {

constexpr int kPairCols = 256;  // columns of a tile, one lane each

// index of D2Q9 direction i among the directions with the same c_y
__host__ __device__ constexpr int ypos(int i) {
  int n = 0;
  for (int k = 0; k < i; ++k)
    if (D2Q9::c(k, 1) == D2Q9::c(i, 1)) ++n;
  return n;
}

// ring slot of direction i for A row p: rows with c_y = -1 keep 1 row, c_y = 0 keep 2,
// c_y = +1 keep 3
__device__ __forceinline__ int ring_slot(int i, int p) {
  const int cy = D2Q9::c(i, 1);
  if (cy < 0) return ypos(i);
  if (cy == 0) return 3 + 3 * ((p + 2) & 1) + ypos(i);
  return 9 + 3 * ((p + 3) % 3) + ypos(i);
}

template <int COLL, int ST, int FM>
__global__ void __launch_bounds__(kPairCols)
k_step_pair(const KArgs a, const typename Codec<ST>::T* __restrict__ A,
            typename Codec<ST>::T* __restrict__ B, int nxt, int Ly) {
  using L = D2Q9;
  using C = Codec<ST>;
  using T = typename C::T;
  constexpr int Q = L::Q, W = kPairCols + 2;
  constexpr uint32_t TS = sizeof(T);
  __shared__ T R[18][W];  // the rings, by ring_slot; ring column c is node column x0 - 1 + c
  const int nx = a.nx, ny = a.ny;
  // Workgroup b runs on XCD b mod 8: give each XCD a contiguous range of tiles, so that the
  // tiles side by side in a row share that XCD's L2 for the cache lines at their edges (the
  // x-shifted loads and the halo columns), instead of two XCDs both fetching them.
  int u = (int)blockIdx.x;
  const int full = (int)gridDim.x & ~7;
  if (u < full) u = (u & 7) * (full >> 3) + (u >> 3);
  const int xt = u % nxt, yc = u / nxt;
  const int x0 = xt * kPairCols;
  const int w = min(kPairCols, nx - x0);  // columns of this tile
  const int y0 = yc * Ly, y1 = min(y0 + Ly, ny);
  const int lane = (int)threadIdx.x;
  const bool wave0 = __builtin_amdgcn_readfirstlane(lane >> 6) == 0;
  const size_t plane = (size_t)a.N * TS;
  const char* Ab = reinterpret_cast<const char*>(A);
  char* Bb = reinterpret_cast<char*>(B);
  // Every lane runs the same loads and stores, so that the compiler's count of outstanding
  // memory operations holds on every path (a branch around a load or store would make it
  // wait for the stores of step B before the next row's loads): lanes past the end of a
  // partial tile repeat its last column (the same values to the same addresses), and every
  // wave loads a halo column (lane 0 the left one, the others the right one), which only
  // the first wave collides.
  const int lc = min(lane, w - 1);
  const int xo = x0 + lc, col = lc + 1;
  const int xh = lane == 0 ? (x0 - 1 + nx) % nx : (x0 + w) % nx;
  const int ch = lane == 0 ? 0 : w + 1;
  // wrap of a row index in [-2 ny, 2 ny) (uniform: the scalar unit)
  auto wrapy = [&](int y) {
    y = y < 0 ? y + ny : y;
    y = y < 0 ? y + ny : y;
    y = y >= ny ? y - ny : y;
    return y >= ny ? y - ny : y;
  };
  const char* ffb = reinterpret_cast<const char*>(a.ff);

  // step A's loads at (x, row p): slot i from (x - c_x, p - c_y), periodic on both axes
  // (the row part in the scalar offset), and the force of the node
  auto issue1 = [&](int x, int p, T (&t)[Q], float (&F)[2]) {
    const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int cx = L::c(i, 0);
      const int xs = cx > 0 ? xm : (cx < 0 ? xp : x);
      const ptrdiff_t row = (ptrdiff_t)nx * wrapy(p - L::c(i, 1)) * (ptrdiff_t)TS;
      t[i] = ld_pl<T, 1>(Ab, plane, i, row, (uint32_t)xs * TS);
    }
    if (FM == kForceField) {
      const uint32_t n = (uint32_t)(x + nx * wrapy(p));
      F[0] = ld_at<float>(ffb, n * 4u);
      F[1] = ld_at<float>(ffb + (size_t)a.N * sizeof(float), n * 4u);
    } else {
      F[0] = FM == kForceUniform ? a.fx : 0.f;
      F[1] = FM == kForceUniform ? a.fy : 0.f;
    }
  };
  // step A at row p from the loaded values: collide, encode into the rings at column c
  auto finish1 = [&](int p, int c, const T (&t)[Q], const float (&F2)[2]) {
    float h[Q], d[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) h[i] = C::ld(t[i]) + (C::shifted ? 0.0f : -(float)L::w(i));
    float F[3] = {F2[0], F2[1], 0.f};
    collide_node<L, COLL, FM != kForceNone>(h, d, F, a.r0, a.r1, a.r2, a.r3);
#pragma unroll
    for (int i = 0; i < Q; ++i)
      R[ring_slot(i, p)][c] = C::st(C::shifted ? d[i] : d[i] + (float)L::w(i));
  };
  T tA[Q], tH[Q];
  float FA[2], FH[2], Fm2[2], Fm1[2];  // forces of the own column at rows r, r - 2, r - 1
  auto issue = [&](int p) {
    issue1(xo, p, tA, FA);
    issue1(xh, p, tH, FH);
  };
  auto finish = [&](int p) {
    finish1(p, col, tA, FA);
    if (wave0) finish1(p, ch, tH, FH);
    Fm2[0] = Fm1[0];
    Fm2[1] = Fm1[1];
    Fm1[0] = FA[0];
    Fm1[1] = FA[1];
  };
  // step B at (x, row q): direction i from A row q - c_y,i, ring column col - c_x,i
  auto stepB = [&](int q) {
    float h[Q], d[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i)
      h[i] = C::ld(R[ring_slot(i, q - L::c(i, 1))][col - L::c(i, 0)]) +
             (C::shifted ? 0.0f : -(float)L::w(i));
    float F[3] = {Fm2[0], Fm2[1], 0.f};
    collide_node<L, COLL, FM != kForceNone>(h, d, F, a.r0, a.r1, a.r2, a.r3);
    const uint32_t off = (uint32_t)(xo + nx * q) * TS;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const float v = C::shifted ? d[i] : d[i] + (float)L::w(i);
      // ordinary stores: the L2 acknowledges them early, and the next row's loads, whose
      // wait also waits for every earlier memory operation, follow closely
      st_pl<0, 1>(Bb, plane, i, 0, off, C::st(v));
    }
  };

  // rows y0 - 1 .. y0 + 1: step A only (y1 >= y0 + 1)
  Fm1[0] = Fm1[1] = 0.f;
  issue(y0 - 1);
  finish(y0 - 1);
  issue(y0);
  finish(y0);
  issue(y0 + 1);
  finish(y0 + 1);
  __syncthreads();
  // rows r = y0 + 2 .. y1: the loads of row r fly while step B runs on row r - 2, then
  // step A of row r.  The loads, the stores and the wait for the loads share one straight
  // block, so the wait is for the loads alone; the scheduling barriers keep step A's
  // arithmetic from moving into step B's (two working sets at once).
  for (int r = y0 + 2; r <= y1; ++r) {
    issue(r);
    __builtin_amdgcn_sched_barrier(0);
    stepB(r - 2);
    __builtin_amdgcn_sched_barrier(0);
    __syncthreads();  // step B has read the ring slots step A of row r overwrites
    finish(r);
    __syncthreads();
  }
  stepB(y1 - 1);
}

// Rows per y-chunk: the height that makes the grid one wave of resident workgroups (CUs x
// workgroups per CU), capped at kPairRowsMax, odd from 3 rows up.  Measured on MI355X (BGK,
// fp32 / fp16, us per step against the height):
//   * small states (512^2: 1 row, 3.7 us against 7.9 at 9; 1024^2: 2 rows, 8.5 / 7.7 us
//     against 9.7 / 10.2 at 9): the sweep's row latency sets the time, so short chunks win
//     although step A then repeats 2 halo rows per chunk (the state sits in the caches);
//   * large states (8192^2: 9 rows, 497 / 281 us against 551 / 306 at one wave of 129-row
//     chunks; 4096^2 within 2% from 5 to 65 rows): many waves of short chunks keep the rows
//     in flight close together, as a one-pass kernel does, which outweighs the 2/9 halo;
//   * chunks a power of two of rows apart, swept in step, meet the same memory channels at
//     once (a quarter of the bandwidth lost at 32 rows with one wave): hence odd heights.
constexpr int kPairRowsMax = 9;

inline int pair_chunk_rows(int nxt, int ny, int slots) {
  int Ly = (int)(((int64_t)ny * nxt + slots - 1) / slots);
  if (Ly >= kPairRowsMax) return kPairRowsMax;
  if (Ly >= 3) Ly |= 1;
  return Ly < 1 ? 1 : Ly;
}

// workgroup slots of kernel k on the current device (cached per device; host threads may
// race to fill an entry, with the same value)
template <class K>
int pair_slots(K k) {
  static std::atomic<int> cache[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 2048;
  const int known = cache[dev].load(std::memory_order_relaxed);
  if (known > 0) return known;
  int cus = 0, nb = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, kPairCols, 0) != hipSuccess ||
      cus <= 0 || nb <= 0)
    return 2048;
  cache[dev].store(cus * nb, std::memory_order_relaxed);
  return cus * nb;
}

template <int COLL, int ST, int FM>
int launch_step_pair_f(const KArgs& a, const void* src, void* dst, hipStream_t s) {
  using T = typename Codec<ST>::T;
  auto k = k_step_pair<COLL, ST, FM>;
  const int nxt = (a.nx + kPairCols - 1) / kPairCols;
  const int Ly = pair_chunk_rows(nxt, a.ny, pair_slots(k));
  const int nyc = (a.ny + Ly - 1) / Ly;
  const int64_t wgs = (int64_t)nxt * nyc;
  if (wgs > INT32_MAX) return PYLBM_ERANGE;
  const dim3 grid((unsigned)wgs), block(kPairCols);
  k<<<grid, block, 0, s>>>(a, (const T*)src, (T*)dst, nxt, Ly);
  return (int)hipGetLastError();
}

// Sweep-stream probe: the memory access pattern of k_step_pair without its arithmetic.  The
// same tiles, chunks of rows (so the same grid), XCD
// order, loads one row ahead and ordinary stores; each row is pulled from src (slot i from
// (x - c_x, y - c_y), periodic) and stored to dst, so dst becomes one streaming step of
// src.  No halo rows: the rate counts the state bytes of a launch.
template <int ST>
__global__ void __launch_bounds__(kPairCols)
k_sweep_probe(const KArgs a, const typename Codec<ST>::T* __restrict__ A,
              typename Codec<ST>::T* __restrict__ B, int nxt, int Ly) {
  using L = D2Q9;
  using T = typename Codec<ST>::T;
  constexpr int Q = L::Q;
  constexpr uint32_t TS = sizeof(T);
  const int nx = a.nx, ny = a.ny;
  int u = (int)blockIdx.x;
  const int full = (int)gridDim.x & ~7;
  if (u < full) u = (u & 7) * (full >> 3) + (u >> 3);
  const int xt = u % nxt, yc = u / nxt;
  const int x0 = xt * kPairCols;
  const int w = min(kPairCols, nx - x0);
  const int y0 = yc * Ly, y1 = min(y0 + Ly, ny);
  const int x = x0 + min((int)threadIdx.x, w - 1);
  const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
  const size_t plane = (size_t)a.N * TS;
  const char* Ab = reinterpret_cast<const char*>(A);
  char* Bb = reinterpret_cast<char*>(B);
  auto wrapy = [&](int y) { return y < 0 ? y + ny : (y >= ny ? y - ny : y); };
  T t[Q];
  auto issue = [&](int p) {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int cx = L::c(i, 0);
      const int xs = cx > 0 ? xm : (cx < 0 ? xp : x);
      const ptrdiff_t row = (ptrdiff_t)nx * wrapy(p - L::c(i, 1)) * (ptrdiff_t)TS;
      t[i] = ld_pl<T, 1>(Ab, plane, i, row, (uint32_t)xs * TS);
    }
  };
  auto store = [&](int q, const T (&v)[Q]) {
    const uint32_t off = (uint32_t)(x + nx * q) * TS;
#pragma unroll
    for (int i = 0; i < Q; ++i) st_pl<0, 1>(Bb, plane, i, 0, off, v[i]);
  };
  issue(y0);
  for (int r = y0; r + 1 < y1; ++r) {
    T v[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) v[i] = t[i];
    issue(r + 1);
    __builtin_amdgcn_sched_barrier(0);
    store(r, v);
  }
  store(y1 - 1, t);
}

// One sweep-stream launch over a periodic nx x ny D2Q9 state of storage ST (the caller checks
// the arguments; state under 4 GiB), chunked as k_step_pair<BGK, ST> is.
template <int ST>
int launch_sweep_probe_t(const KArgs& a, const void* src, void* dst, hipStream_t s) {
  using T = typename Codec<ST>::T;
  const int nxt = (a.nx + kPairCols - 1) / kPairCols;
  const int Ly = pair_chunk_rows(nxt, a.ny, pair_slots(k_step_pair<PYLBM_BGK, ST, kForceNone>));
  const int nyc = (a.ny + Ly - 1) / Ly;
  const int64_t wgs = (int64_t)nxt * nyc;
  if (wgs > INT32_MAX) return PYLBM_ERANGE;
  k_sweep_probe<ST><<<dim3((unsigned)wgs), dim3(kPairCols), 0, s>>>(a, (const T*)src, (T*)dst,
                                                                     nxt, Ly);
  return (int)hipGetLastError();
}

template <int COLL, int ST>
int launch_step_pair_t(const KArgs& a, int fmode, const void* src, void* dst, hipStream_t s) {
  if (fmode == kForceField) return launch_step_pair_f<COLL, ST, kForceField>(a, src, dst, s);
  if (fmode == kForceUniform) return launch_step_pair_f<COLL, ST, kForceUniform>(a, src, dst, s);
  return launch_step_pair_f<COLL, ST, kForceNone>(a, src, dst, s);
}

template <int COLL>
int launch_step_pair_s(const KArgs& a, int storage, int fmode, const void* src, void* dst,
                       hipStream_t s) {
  switch (storage) {
    case PYLBM_F32: return launch_step_pair_t<COLL, PYLBM_F32>(a, fmode, src, dst, s);
    case PYLBM_F16: return launch_step_pair_t<COLL, PYLBM_F16>(a, fmode, src, dst, s);
    case PYLBM_BF16: return launch_step_pair_t<COLL, PYLBM_BF16>(a, fmode, src, dst, s);
    case PYLBM_F32S: return launch_step_pair_t<COLL, PYLBM_F32S>(a, fmode, src, dst, s);
  }
  return PYLBM_EINVAL;
}

// Two steps src -> dst on a D2Q9 box periodic on both axes whose state is under 4 GiB (one
// descriptor per array; the caller checks).
inline int launch_step_pair(const KArgs& a, int coll, int storage, int fmode, const void* src,
                            void* dst, hipStream_t s) {
  switch (coll) {
    case PYLBM_BGK: return launch_step_pair_s<PYLBM_BGK>(a, storage, fmode, src, dst, s);
    case PYLBM_TRT: return launch_step_pair_s<PYLBM_TRT>(a, storage, fmode, src, dst, s);
    case PYLBM_MRT: return launch_step_pair_s<PYLBM_MRT>(a, storage, fmode, src, dst, s);
  }
  return PYLBM_EINVAL;
}

}  // namespace plbm
