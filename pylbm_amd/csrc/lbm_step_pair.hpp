// Two time steps per launch (temporal blocking) of the two-buffer pull, on D2Q9 boxes
// (each axis periodic or walled with half-way bounce-back, Ladd's lid on the top wall; no
// obstacles, inlet or outlet): src -> dst over two steps, the intermediate state kept on chip.
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
//     three newest (18 values; with walls the c_y = -1 ones of the two newest), for the
//     tile's columns and one halo column on each side, which the first wave computes;
//   * step A also runs on the halo rows y0 - 1 and y1 (1 + 2 / Ly of the reads);
//   * step A's loads of row r are issued before step B of row r - 2 runs, so that they fly
//     while it computes; the y-chunks are short (pair_chunk_rows), and each XCD takes a
//     contiguous range of tiles.
// Per node and two steps: src read once (plus the halo), dst written once, half the bytes of
// two single steps.  Every value is the single steps' value: the same gather, decode and
// collision on the same encoded populations.
#pragma once

#include <atomic>
#include <type_traits>

#include "lbm_step.hpp"

namespace plbm {

constexpr int kPairCols = 256;  // columns of a tile, one lane each
constexpr int kXPos[3] = {1, 5, 8};  // the D2Q9 directions with c_x > 0

// index of D2Q9 direction i among the directions with the same c_y
__host__ __device__ constexpr int ypos(int i) {
  int n = 0;
  for (int k = 0; k < i; ++k)
    if (D2Q9::c(k, 1) == D2Q9::c(i, 1)) ++n;
  return n;
}

// ring slot of direction i for A row p: the components with c_y = -1 keep RM rows (1 on
// periodic boxes; 2 with walls, where an x-wall link of row q and the bottom wall's
// bounce-back read the own node's component of row q after row q + 1 is in), c_y = 0 keep
// 2 rows, c_y = +1 keep 3
template <int RM>
__device__ __forceinline__ int ring_slot(int i, int p) {
  const int cy = D2Q9::c(i, 1);
  if (cy < 0) return (RM == 1 ? 0 : 3 * ((p + 2) & 1)) + ypos(i);
  if (cy == 0) return 3 * RM + 3 * ((p + 2) & 1) + ypos(i);
  return 3 * RM + 6 + 3 * ((p + 3) % 3) + ypos(i);
}

// Ladd's moving-wall term of link i when it crosses the lid, as the one-step gather forms it
template <class L>
__device__ __forceinline__ float lid_term(int i, float ulid) {
  return (float)(6.0 * L::w(i) * L::c(i, 0)) * ulid;
}

// WALLS: half-way bounce-back on the non-periodic axes (a.per) and Ladd's lid on the top
// wall; false: periodic on both axes (no wall logic compiled in)
template <int COLL, int ST, int FM, bool WALLS>
__global__ void __launch_bounds__(kPairCols)
k_step_pair(const KArgs a, const typename Codec<ST>::T* __restrict__ A,
            typename Codec<ST>::T* __restrict__ B, int nxt, int Ly) {
  using L = D2Q9;
  using C = Codec<ST>;
  using T = typename C::T;
  constexpr int Q = L::Q, W = kPairCols + 2;
  constexpr int RM = WALLS ? 2 : 1, NR = 3 * RM + 15;
  constexpr uint32_t TS = sizeof(T);
  __shared__ T R[NR][W];  // the rings, by ring_slot; ring column c is node column x0 - 1 + c
  const int nx = a.nx, ny = a.ny;
  // 1 where an axis has walls.  The wall tests below combine such ints with & and |, never
  // with && or ||: a short-circuit test becomes a branch between the loads, and a branch
  // makes the compiler's count of outstanding memory operations cover step B's stores.
  const int wlx = WALLS ? (~a.per & 1) : 0, wly = WALLS ? ((~a.per >> 1) & 1) : 0;
  // 1 when row ys lies behind a y wall (ys in [-2, ny + 1])
  auto ywall = [&](int ys) { return wly & (int)((unsigned)ys >= (unsigned)ny); };
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
  // partial tile repeat its last column (the same values to the same addresses), and the
  // waves that load the halo columns (lane 0 the left one, the others the right one) do so
  // in every row, though only the first wave collides them.
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

  // The kernel body in three forms, chosen per workgroup (so every form keeps its loads and
  // stores straight):
  //   MODE 0, interior tiles (neither the tile nor its halo columns touch x = 0 or nx - 1):
  //     no column wraps, so every load of a node has the same vector offset (its column)
  //     and the x shift of a link goes into the scalar offset with the slot and the row;
  //   MODE 1, the first and last tile of a row, x periodic: per-lane wrapped columns;
  //   MODE 2, the first and last tile of a row, x walls: as MODE 1, plus the own node's
  //     slots for the x-wall links.
  // A y wall is uniform per row: its links select slot, row and shift in scalar registers in
  // step A, and in step B they read the rings of the rows behind the wall (-1 and ny), which
  // hold the bounced-back values (bounce(), below), so that step B has no wall test except
  // the per-lane x-wall one of MODE 2.
  auto body = [&](auto mode_tag, auto halo_tag) {
    constexpr int MODE = decltype(mode_tag)::value;
    // HALO: this wave loads the halo columns (see dispatch below)
    constexpr bool HALO = decltype(halo_tag)::value;
    // Step A's loads at (x, row p): slot i from (x - c_x, p - c_y), periodic, and the force
    // of the node.  With walls, a link whose source lies behind a wall reads the own node's
    // slot opp(i) instead (half-way bounce-back): a y-wall link for the whole row (slot,
    // row and shift in scalar registers); in MODE 2 every lane also loads its own slots
    // opp(i) of the links with c_x != 0, which the lanes on an x wall take in place of the
    // gathered ones (xwall).  No branch between the loads.
    auto issue1 = [&](int x, bool halo, int p, T (&t)[Q], T (&e)[3], float (&F)[2]) {
      if constexpr (MODE == 0) {
        const uint32_t vo = (uint32_t)x * TS;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
          const int cx = L::c(i, 0), cy = L::c(i, 1), ys = p - cy;
          const bool wy = WALLS && ywall(ys) != 0;  // uniform
          // slot, row and shift in one scalar offset (the state is under 4 GiB); the
          // readfirstlane tells the compiler that it is uniform, which it cannot prove
          // through the wall selects (it would loop the load over the lanes' values)
          const uint32_t so = (uint32_t)((size_t)(wy ? L::opp(i) : i) * plane) +
                              (uint32_t)(nx * wrapy(wy ? p : ys)) * TS -
                              (uint32_t)(wy ? 0 : cx) * TS;
          t[i] = ld_pl<T, 1>(Ab, plane, 0, (ptrdiff_t)__builtin_amdgcn_readfirstlane((int)so),
                             vo);
        }
      } else {
        const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
          const int cx = L::c(i, 0), cy = L::c(i, 1), ys = p - cy;
          const int xs = cx > 0 ? xm : (cx < 0 ? xp : x);
          if constexpr (WALLS) {
            const bool wy = ywall(ys) != 0;  // uniform
            const ptrdiff_t row = (ptrdiff_t)nx * (wy ? wrapy(p) : wrapy(ys)) * (ptrdiff_t)TS;
            t[i] = ld_pl<T, 1>(Ab, plane, wy ? L::opp(i) : i, row, (uint32_t)(wy ? x : xs) * TS);
          } else {
            const ptrdiff_t row = (ptrdiff_t)nx * wrapy(ys) * (ptrdiff_t)TS;
            t[i] = ld_pl<T, 1>(Ab, plane, i, row, (uint32_t)xs * TS);
          }
        }
      }
      if constexpr (MODE == 2) {
        // own slots of row p for the x-wall links (xwall() picks them once the loads are in):
        // e[k] is slot opp(i) of pair k's link i = kXPos[k] (c_x > 0) at x = 0, else slot i,
        // which the link opp(i) (c_x < 0) takes at x = nx - 1; the lower slot of the pair in
        // the scalar offset, the plane difference (0 or 2 planes) per lane.  For the halo
        // (halo = true) only x = nx - 1 matters: a halo column at x = 0 is a wrapped one
        // (nx > 1) or the own column itself (nx = 1), whose e holds the other side.
        const ptrdiff_t rowp = (ptrdiff_t)nx * wrapy(p) * (ptrdiff_t)TS;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int ip = kXPos[k], im = L::opp(ip), s0 = ip < im ? ip : im;
          const uint32_t dp = (uint32_t)((size_t)(((!halo && x == 0) ? im : ip) - s0) * plane);
          e[k] = ld_pl<T, 1>(Ab, plane, s0, rowp, dp + (uint32_t)x * TS);
        }
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
    // decode of the value of link i with source row ys: the lid term on the links with
    // c_y = -1, c_x != 0 through the top wall (after any bounce-back, as the one-step gather
    // adds it; the other links add nothing: the decode's + 0 already made -0 into +0)
    auto decode = [&](int i, T v, int ys) {
      float h = C::ld(v) + (C::shifted ? 0.0f : -(float)L::w(i));
      if (WALLS && L::c(i, 1) < 0 && L::c(i, 0) != 0)
        h = h + ((wly & (int)(ys >= ny)) ? lid_term<L>(i, a.ulid) : 0.0f);
      return h;
    };
    // step A at row p from the loaded values: collide, encode into the rings at column c
    auto finish1 = [&](int p, int c, const T (&t)[Q], const float (&F2)[2]) {
      float h[Q], d[Q];
#pragma unroll
      for (int i = 0; i < Q; ++i) h[i] = decode(i, t[i], p - L::c(i, 1));
      float F[3] = {F2[0], F2[1], 0.f};
      collide_node<L, COLL, FM != kForceNone>(h, d, F, a.r0, a.r1, a.r2, a.r3);
#pragma unroll
      for (int i = 0; i < Q; ++i)
        R[ring_slot<RM>(i, p)][c] = C::st(C::shifted ? d[i] : d[i] + (float)L::w(i));
    };
    T tA[Q], tH[Q], eA[3], eH[3];
    float FA[2], FH[2], Fm2[2], Fm1[2];  // forces of the own column at rows r, r - 2, r - 1
    auto issue = [&](int p) {
      issue1(xo, false, p, tA, eA, FA);
      if constexpr (HALO) issue1(xh, true, p, tH, eH, FH);
    };
    // MODE 2: a link that crosses an x wall (c_x > 0 at x = 0, c_x < 0 at x = nx - 1; also
    // at a corner, where the y-wall load holds the same value) takes the own node's slot.
    // For the own column at x = 0 = nx - 1 (nx = 1) the c_x < 0 side comes from the halo
    // loads, which are of the same node there.
    auto xwall = [&](int x, bool halo, T (&t)[Q], const T (&e)[3], const T (&eh)[3]) {
      if constexpr (MODE == 2) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int ip = kXPos[k], im = L::opp(ip);
          if (!halo) t[ip] = x == 0 ? e[k] : t[ip];
          t[im] = x == nx - 1 ? ((!halo && x == 0) ? eh[k] : e[k]) : t[im];
        }
      }
    };
    auto finish = [&](int p) {
      // eH serves the own column only for nx = 1, where every wave loads the halo
      if constexpr (HALO) xwall(xo, false, tA, eA, eH);
      else xwall(xo, false, tA, eA, eA);
      finish1(p, col, tA, FA);
      if constexpr (HALO) {
        xwall(xh, true, tH, eH, eH);
        if constexpr (WALLS) {
          // every wave that loads the halo takes those loads here (only the first wave
          // collides them), so that none is still counted as outstanding when its registers
          // are reused in the next row (which would make the compiler wait there for step
          // B's stores too)
#pragma unroll
          for (int i = 0; i < Q; ++i) asm volatile("" ::"v"(tH[i]));
        }
        if (wave0) finish1(p, ch, tH, FH);
      }
      Fm2[0] = Fm1[0];
      Fm2[1] = Fm1[1];
      Fm1[0] = FA[0];
      Fm1[1] = FA[1];
    };
    // Bounce-back at a y wall for step B: A row pw lies behind the wall next to A row pn,
    // and a link c_y = cyw (toward row pn) from row pw into row pn takes the own node's
    // component opp(i) of A row pn.  Each lane writes, for its own column, the value that
    // the step-B node there reads: ring (i, pw) at column col - c_x = ring (opp(i), pn) at
    // col (the ring columns 0 .. w + 1 cover every read of step B).
    auto bounce = [&](int pn, int pw, int cyw) {
#pragma unroll
      for (int i = 0; i < Q; ++i)
        if (L::c(i, 1) == cyw)
          R[ring_slot<RM>(i, pw)][col - L::c(i, 0)] = R[ring_slot<RM>(L::opp(i), pn)][col];
    };
    // step B at (x, row q): direction i from A row q - c_y,i, ring column col - c_x,i; in
    // MODE 2 an x-wall link takes the own node's component opp(i) of A row q
    auto stepB = [&](int q) {
      float h[Q], d[Q];
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const int cx = L::c(i, 0), cy = L::c(i, 1);
        if constexpr (MODE == 2) {
          const int wx = cx > 0 ? (int)(xo == 0) : (cx < 0 ? (int)(xo == nx - 1) : 0);
          // one LDS read from the selected slot and column
          const int sl = wx ? ring_slot<RM>(L::opp(i), q) : ring_slot<RM>(i, q - cy);
          h[i] = decode(i, R[sl][wx ? col : col - cx], q - cy);
        } else {
          h[i] = decode(i, R[ring_slot<RM>(i, q - cy)][col - cx], q - cy);
        }
      }
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
    if constexpr (WALLS) {
      if (wly & (int)(y0 == 0)) {  // the bottom wall: row -1 (A row 0 is still in the rings)
        bounce(0, -1, 1);
        __syncthreads();
      }
    }
    // rows r = y0 + 2 .. y1: the loads of row r fly while step B runs on row r - 2, then
    // step A of row r.  The loads, the stores and the wait for the loads share one
    // straight block, so the wait is for the loads alone; the scheduling barriers keep step
    // A's arithmetic from moving into step B's (two working sets at once).
    for (int r = y0 + 2; r <= y1; ++r) {
      issue(r);
      __builtin_amdgcn_sched_barrier(0);
      stepB(r - 2);
      __builtin_amdgcn_sched_barrier(0);
      __syncthreads();  // step B has read the ring slots step A of row r overwrites
      finish(r);
      __syncthreads();
    }
    if constexpr (WALLS) {
      if (wly & (int)(y1 == ny)) {  // the top wall: row ny
        bounce(ny - 1, ny, -1);
        __syncthreads();
      }
    }
    stepB(y1 - 1);
  };
  // Only the first wave collides the halo columns, so the other waves run a copy of the
  // body without the halo loads.  The choice is per wave and made once, outside the row
  // loop, so each copy keeps its loads and stores straight; both copies pass the same
  // workgroup barriers.  With x walls and nx = 1 every wave loads the halo: the own column
  // takes its x-wall slots from those loads (the halo column is the own node there).
  using HaloOn = std::true_type;
  using HaloOff = std::false_type;
  if (xt != 0 && x0 + w < nx - 1) {
    if (wave0) body(std::integral_constant<int, 0>{}, HaloOn{});
    else body(std::integral_constant<int, 0>{}, HaloOff{});
  } else if (WALLS && wlx) {
    if (wave0 || nx == 1) body(std::integral_constant<int, 2>{}, HaloOn{});
    else body(std::integral_constant<int, 2>{}, HaloOff{});
  } else {
    if (wave0) body(std::integral_constant<int, 1>{}, HaloOn{});
    else body(std::integral_constant<int, 1>{}, HaloOff{});
  }
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

// workgroup slots of kernel KERNEL on the current device, cached per kernel and device
// (each instantiation of this template has its own cache: kernels of one signature differ in
// their registers and so in their occupancy; host threads may race to fill an entry, with
// the same value)
template <auto KERNEL>
int pair_slots() {
  static std::atomic<int> cache[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 2048;
  const int known = cache[dev].load(std::memory_order_relaxed);
  if (known > 0) return known;
  int cus = 0, nb = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, KERNEL, kPairCols, 0) != hipSuccess ||
      cus <= 0 || nb <= 0)
    return 2048;
  cache[dev].store(cus * nb, std::memory_order_relaxed);
  return cus * nb;
}

template <int COLL, int ST, int FM, bool WALLS>
int launch_step_pair_w(const KArgs& a, const void* src, void* dst, hipStream_t s) {
  using T = typename Codec<ST>::T;
  auto k = k_step_pair<COLL, ST, FM, WALLS>;
  const int nxt = (a.nx + kPairCols - 1) / kPairCols;
  const int Ly = pair_chunk_rows(nxt, a.ny, pair_slots<k_step_pair<COLL, ST, FM, WALLS>>());
  const int nyc = (a.ny + Ly - 1) / Ly;
  const int64_t wgs = (int64_t)nxt * nyc;
  if (wgs > INT32_MAX) return PYLBM_ERANGE;
  const dim3 grid((unsigned)wgs), block(kPairCols);
  k<<<grid, block, 0, s>>>(a, (const T*)src, (T*)dst, nxt, Ly);
  return (int)hipGetLastError();
}

// the wall logic only where an axis has walls
template <int COLL, int ST, int FM>
int launch_step_pair_f(const KArgs& a, const void* src, void* dst, hipStream_t s) {
  if ((a.per & 3) == 3) return launch_step_pair_w<COLL, ST, FM, false>(a, src, dst, s);
  return launch_step_pair_w<COLL, ST, FM, true>(a, src, dst, s);
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
  const int Ly =
      pair_chunk_rows(nxt, a.ny, pair_slots<k_step_pair<PYLBM_BGK, ST, kForceNone, false>>());
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

// Two steps src -> dst on a D2Q9 box (periodic axes or half-way bounce-back walls, the lid)
// whose state is under 4 GiB (one descriptor per array; the caller checks).
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
