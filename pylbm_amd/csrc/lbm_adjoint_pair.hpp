// Two adjoint steps per launch (temporal blocking) on D3Q19 boxes periodic on every axis, from
// moment records, with fp32 adjoint states in two buffers (lam_in -> lam_out).
//
// The adjoint of a rollout moves the adjoint state through memory twice per step (read
// and write, 152 B per node in fp32) and a 16 B moment record; chained single steps are at
// the memory system's rate.  Here the first step's result (mu_A, pushed along its links, is
// the second step's input) never leaves the chip:
//
//   * a workgroup holds kPairRows + 2 rows of full x-rows (nx <= kPairLanes) of one z-chunk
//     and streams through its planes, z0 - 1 .. z1 (z1 = z0 + Lz, or nz);
//   * step A runs on every row (the tile and one halo row on each side) of plane p and
//     keeps mu_A in LDS; step B runs on the tile's rows of plane p - 1 once step A has done
//     plane p, gathering lam_B,k(x) = mu_A,k(x + c_k) from LDS, and pushes its mu_B to
//     lam_out in memory;
//   * the loads of plane p (and the record of step B's next plane) are issued before step B
//     of plane p - 2 runs, and the wait for them covers no store (every lane of a wave runs
//     the same loads and stores);
//   * LDS keeps, per column, only the components a later plane still needs: those with
//     c_z = +1 of the newest A plane (B of the plane below reads them), c_z = 0 of the two
//     newest, c_z = -1 of the three newest: 38 fp32 per column, 117 KB per workgroup;
//   * workgroup b takes z-chunk b % nzc and y-tile b / nzc: with 8 z-chunks each XCD sweeps
//     one chunk, so the halo rows two neighbouring tiles both read meet in that XCD's L2.
//
// Per node and two steps: lam_in read for the tile and its halo rows and planes, two
// records, lam_out written once.  Every lam_out slot is written exactly once (a push is a
// permutation), and the values equal two single record steps bit for bit: the same
// adjoint_core on the same fp32 inputs (LDS holds mu_A exactly as a stored fp32 lam would).
#pragma once

#include "lbm_adjoint_kernels.hpp"

namespace plbm {
namespace adj {

constexpr int kPairRows = 4;     // interior rows of a tile
constexpr int kPairLanes = 128;  // lanes per row: rows of up to 128 nodes
constexpr int kPairBlock = kPairLanes * (kPairRows + 2);

// index of direction i among the D3Q19 directions with the same c_z
__host__ __device__ constexpr int zpos(int i) {
  int n = 0;
  for (int k = 0; k < i; ++k)
    if (D3Q19::c(k, 2) == D3Q19::c(i, 2)) ++n;
  return n;
}

// SP: cache policy of the lam_out stores (0 temporal, 2 non-temporal)
template <int COLL, int FM, int SP>
__global__ void __launch_bounds__(kPairBlock)
k_adjoint_pair(const KArgs a, const float* __restrict__ recA, const float* __restrict__ recB,
               const float* __restrict__ lam_in, float* __restrict__ lam_out, int nzc, int Lz) {
  using L = D3Q19;
  constexpr int Q = L::Q, R = kPairRows + 2, W = kPairLanes;
  __shared__ float sp1[5][R][W];     // c_z = +1 components of the newest A plane
  __shared__ float s0[2][9][R][W];   // c_z = 0 components of the two newest
  __shared__ float sm1[3][5][R][W];  // c_z = -1 components of the three newest
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int zc = (int)blockIdx.x % nzc, yt = (int)blockIdx.x / nzc;
  const int z0 = zc * Lz;
  const int z1 = min(z0 + Lz, nz);  // planes [z0, z1) are this workgroup's
  if (z0 >= z1) return;             // an empty chunk (nz < nzc * Lz): the whole workgroup
  // Every lane of a wave runs the same loads and stores, so that the compiler's count of
  // outstanding memory operations holds on every path (a branch around one makes a wait
  // for a load also wait for the stores issued before it): lanes past the end of a row
  // repeat its last node (the same values to the same addresses).  Whether a wave runs
  // step B is uniform per wave (its row), and the two kinds of wave run separate loops.
  const int x = min((int)threadIdx.x % W, nx - 1);
  const int r = __builtin_amdgcn_readfirstlane((int)threadIdx.x / W);  // uniform per wave
  const int y0 = yt * kPairRows;
  const int y = (y0 - 1 + r + ny) % ny;  // the row of lane row r (halo rows wrap)
  const bool brow = r >= 1 && r <= kPairRows && y0 + r - 1 < ny;  // uniform per wave
  const size_t N = (size_t)a.N;
  const char* lib = reinterpret_cast<const char*>(lam_in);
  char* lob = reinterpret_cast<char*>(lam_out);
  const size_t lplane = N * 4u;
  const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
  const float F[3] = {FM == 1 ? a.fx : 0.f, FM == 1 ? a.fy : 0.f, FM == 1 ? a.fz : 0.f};

  auto node = [&](int yy, int zz) { return (uint32_t)(x + nx * (yy + ny * zz)); };
  auto wrapz = [&](int p) { return (p % nz + nz) % nz; };
  float lamA[Q], moA[4], moB[4];
  // step A's loads of plane p (any integer; wrapped): lam_in at the own node and its record
  auto issueA = [&](int p) {
    const uint32_t off = node(y, wrapz(p)) * 4u;
#pragma unroll
    for (int i = 0; i < Q; ++i) lamA[i] = ld_pl<float, 1>(lib, lplane, i, 0, off);
    ld_rec_nt(recA, node(y, wrapz(p)), moA);
  };
  // step A at plane p: mu_A into the rings
  auto finishA = [&](int p) {
    float mu[Q], dFn[3], u[3] = {moA[1], moA[2], moA[3]};
    const float rho = moA[0], inv = __builtin_amdgcn_rcpf(rho);
    adjoint_core<L, COLL, FM != 0>(rho, inv, u, F, lamA, a.r0, a.r1, mu, dFn);
    const int s2 = (p + 2) & 1, s3 = (p + 3) % 3;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      if (L::c(i, 2) > 0) sp1[zpos(i)][r][x] = mu[i];
      else if (L::c(i, 2) == 0) s0[s2][zpos(i)][r][x] = mu[i];
      else sm1[s3][zpos(i)][r][x] = mu[i];
    }
  };
  // step B at plane q (A has done q + 1), from the record in mo
  auto stepB = [&](int q, const float (&mo)[4]) {
    // lam_B,k(x) = mu_A,k(x + c_k): plane q + c_z,k, row r + c_y,k, column x + c_x,k
    float lam[Q];
    const int t2 = (q + 2) & 1, t3 = (q + 2) % 3;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      const int cx = L::c(k, 0), cy = L::c(k, 1), cz = L::c(k, 2);
      const int xs = cx > 0 ? xp : (cx < 0 ? xm : x);
      const int rs = r + cy;
      if (cz > 0) lam[k] = sp1[zpos(k)][rs][xs];
      else if (cz == 0) lam[k] = s0[t2][zpos(k)][rs][xs];
      else lam[k] = sm1[t3][zpos(k)][rs][xs];
    }
    float mu[Q], dFn[3], u[3] = {mo[1], mo[2], mo[3]};
    const float rho = mo[0], inv = __builtin_amdgcn_rcpf(rho);
    adjoint_core<L, COLL, FM != 0>(rho, inv, u, F, lam, a.r0, a.r1, mu, dFn);
    // push mu_B,k to (k, x - c_k), wrapped on every axis
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      const int cx = L::c(k, 0), cy = L::c(k, 1), cz = L::c(k, 2);
      const int xt = cx > 0 ? xm : (cx < 0 ? xp : x);
      const int yt2 = (y - cy + ny) % ny, zt = (q - cz + nz) % nz;
      const uint32_t off = (uint32_t)(xt + nx * (yt2 + ny * zt)) * 4u;
      st_pl<SP, 1>(lob, lplane, k, 0, off, mu[k]);
    }
  };

  // Planes z0 - 1 .. z0 + 1: step A only (z1 >= z0 + 1).  A step-B wave loads the record of
  // its first plane before two planes of step-A loads, as many memory operations as the
  // loop issues after each record it loads (the record of plane q is loaded one iteration
  // before step B of plane q uses it, ahead of the iteration's step-A loads): the wait for
  // it then never covers a store.
  issueA(z0 - 1);
  finishA(z0 - 1);
  if (brow) ld_rec_nt(recB, node(y, z0), moB);
  issueA(z0);
  finishA(z0);
  issueA(z0 + 1);
  finishA(z0 + 1);
  __syncthreads();
  if (brow) {
    // planes p = z0 + 2 .. z1: the loads of plane p (and the record of step B's next
    // plane) fly while step B runs on plane p - 2; then step A of plane p
    for (int p = z0 + 2; p <= z1; ++p) {
      float mo[4] = {moB[0], moB[1], moB[2], moB[3]};
      ld_rec_nt(recB, node(y, p - 1), moB);
      issueA(p);
      __builtin_amdgcn_sched_barrier(0);
      stepB(p - 2, mo);
      __builtin_amdgcn_sched_barrier(0);
      __syncthreads();  // step B has read the ring slots step A of plane p overwrites
      finishA(p);
      __syncthreads();
    }
    stepB(z1 - 1, moB);
  } else {
    for (int p = z0 + 2; p <= z1; ++p) {
      issueA(p);
      __builtin_amdgcn_sched_barrier(0);
      __syncthreads();
      finishA(p);
      __syncthreads();
    }
  }
}

// Host side (the caller checks that the kernel applies: adjoint_pair_applies).
template <int COLL>
int launch_pair_c(const KArgs& a, const void* recA, const void* recB, const void* lin,
                  void* lout, hipStream_t s) {
  const int nzc = a.nz < 8 ? a.nz : 8;
  const int Lz = (a.nz + nzc - 1) / nzc;
  const int nyt = (a.ny + kPairRows - 1) / kPairRows;
  const dim3 grid((unsigned)(nyt * nzc)), block(kPairBlock);
  const bool uniform = a.fx != 0.0f || a.fy != 0.0f || a.fz != 0.0f;
  const float* ra = (const float*)recA;
  const float* rb = (const float*)recB;
  const float* li = (const float*)lin;
  float* lo = (float*)lout;
  if (uniform)
    k_adjoint_pair<COLL, 1, 0><<<grid, block, 0, s>>>(a, ra, rb, li, lo, nzc, Lz);
  else
    k_adjoint_pair<COLL, 0, 0><<<grid, block, 0, s>>>(a, ra, rb, li, lo, nzc, Lz);
  return (int)hipGetLastError();
}

}  // namespace adj
}  // namespace plbm
