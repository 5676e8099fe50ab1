// Kernels on either side of the fused step: canonical-state encode / decode, device-side
// equilibrium initialisation, float64 observables and reductions, and the streaming copy
// that measures the bandwidth ceiling.
//
// The float64 arithmetic here is written with contraction off and in the reference's
// evaluation order, so that results equal the reference's host computations on the same
// populations bit for bit:
//   encode  : pylbm/backends/_fused.py:83-90 (f - w_i in float64 -> float32 -> storage)
//   decode  : pylbm/backends/_fused.py:95-103 (exact widening, + w_i, AA unswap)
//   init    : pylbm/lattice.py:112-119, pylbm/problem.py:108-115
//   moments : pylbm/core.py:151-169, 255-266
#include "lbm_aux.hpp"

#include <mutex>

#pragma clang fp contract(off)

namespace plbm {

namespace {

template <class L, int ST>
__global__ void __launch_bounds__(kBlock)
k_decode(const typename Codec<ST>::T* __restrict__ A, double* __restrict__ out, int N, int aa) {
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= N) return;
#pragma unroll
  for (int i = 0; i < L::Q; ++i) {
    const int slot = aa ? L::opp(i) : i;
    double v = (double)Codec<ST>::ld(A[(size_t)slot * N + n]);
    if (Codec<ST>::shifted) v = v + L::w(i);
    out[(size_t)i * N + n] = v;
  }
}

template <class L, int ST>
__device__ __forceinline__ void encode_node(typename Codec<ST>::T* A, const double (&f)[L::Q],
                                            size_t N, int n, int aa) {
#pragma unroll
  for (int i = 0; i < L::Q; ++i) {
    double v = f[i];
    if (Codec<ST>::shifted) v = v - L::w(i);
    const int slot = aa ? L::opp(i) : i;
    A[(size_t)slot * N + n] = Codec<ST>::st((float)v);
  }
}

template <class L, int ST>
__global__ void __launch_bounds__(kBlock)
k_encode(const double* __restrict__ in, typename Codec<ST>::T* __restrict__ A, int N, int aa) {
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= N) return;
  double f[L::Q];
#pragma unroll
  for (int i = 0; i < L::Q; ++i) f[i] = in[(size_t)i * N + n];
  encode_node<L, ST>(A, f, (size_t)N, n, aa);
}

template <class L, int ST>
__global__ void __launch_bounds__(kBlock)
k_init_eq(const double* __restrict__ rho, const double* __restrict__ u,
          typename Codec<ST>::T* __restrict__ A, int N, int aa) {
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= N) return;
  const double r = rho ? rho[n] : 1.0;
  double uu[3] = {0.0, 0.0, 0.0};
  if (u) {
#pragma unroll
    for (int a = 0; a < L::D; ++a) uu[a] = u[(size_t)a * N + n];
  }
  double usq = 0.0;
#pragma unroll
  for (int a = 0; a < L::D; ++a) usq = usq + uu[a] * uu[a];
  double f[L::Q];
#pragma unroll
  for (int i = 0; i < L::Q; ++i) {
    double cu = 0.0;
#pragma unroll
    for (int a = 0; a < L::D; ++a) cu = cu + (double)L::c(i, a) * uu[a];
    f[i] = L::w(i) * r * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq);
  }
  encode_node<L, ST>(A, f, (size_t)N, n, aa);
}

// rho and u of the post-collision state at node n, in the reference's summation order.
template <class L, int ST>
__device__ __forceinline__ void node_moments(const typename Codec<ST>::T* A, size_t N, int n,
                                             int aa, const double F[3], bool has_force,
                                             const bool Fnz[3], double& rho, double (&u)[3]) {
  double f[L::Q];
#pragma unroll
  for (int i = 0; i < L::Q; ++i) {
    const int slot = aa ? L::opp(i) : i;
    double v = (double)Codec<ST>::ld(A[(size_t)slot * N + n]);
    if (Codec<ST>::shifted) v = v + L::w(i);
    f[i] = v;
  }
  rho = f[0];
#pragma unroll
  for (int i = 1; i < L::Q; ++i) rho = rho + f[i];
#pragma unroll
  for (int a = 0; a < L::D; ++a) {
    double j = 0.0;
    bool first = true;
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
      const int c = L::c(i, a);
      if (c == 0) continue;
      const double t = (c == 1) ? f[i] : -f[i];
      j = first ? t : j + t;
      first = false;
    }
    if (Fnz[a]) j = j + 0.5 * F[a];
    u[a] = j / rho;
  }
  if (has_force) {
#pragma unroll
    for (int a = 0; a < L::D; ++a) u[a] = (u[a] * rho - F[a]) / rho;
  }
  if (L::D == 2) u[2] = 0.0;
}

struct MomArgs {
  int N;
  int aa;
  int has_force;
  double F[3];
  int Fnz[3];
  const float* ff;  // per-cell force field (d, N) or nullptr
};

template <class L, int ST>
__device__ __forceinline__ void moments_at(const typename Codec<ST>::T* A, const MomArgs& m,
                                           int n, double& rho, double (&u)[3]) {
  double F[3] = {m.F[0], m.F[1], m.F[2]};
  bool Fnz[3] = {m.Fnz[0] != 0, m.Fnz[1] != 0, m.Fnz[2] != 0};
  if (m.ff) {
#pragma unroll
    for (int a = 0; a < L::D; ++a) {
      F[a] = (double)m.ff[(size_t)a * m.N + n];
      Fnz[a] = true;
    }
  }
  node_moments<L, ST>(A, (size_t)m.N, n, m.aa, F, m.has_force != 0, Fnz, rho, u);
}

template <class L, int ST>
__global__ void __launch_bounds__(kBlock)
k_macro(const typename Codec<ST>::T* __restrict__ A, MomArgs m, double* __restrict__ rho_out,
        double* __restrict__ u_out) {
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= m.N) return;
  double rho, u[3];
  moments_at<L, ST>(A, m, n, rho, u);
  if (rho_out) rho_out[n] = rho;
  if (u_out) {
#pragma unroll
    for (int a = 0; a < L::D; ++a) u_out[(size_t)a * m.N + n] = u[a];
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
// max that keeps a NaN from either side (fmax would drop it), so a NaN or an Inf - Inf
// difference makes MAXABS_DIFF NaN instead of hiding it
__device__ __forceinline__ double max_nan(double a, double b) {
  return (a > b || a != a) ? a : b;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max_nan(v, __shfl_down(v, off, 64));
  return v;
}

// Sum (or max) over the workgroup in a fixed order; thread 0 returns the result.
template <bool MAX>
__device__ double block_reduce(double v) {
  __shared__ double part[kBlock / 64];
  v = MAX ? wave_max(v) : wave_sum(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) part[wv] = v;
  __syncthreads();
  double t = part[0];
  if (threadIdx.x == 0)
    for (int k = 1; k < kBlock / 64; ++k) t = MAX ? max_nan(t, part[k]) : t + part[k];
  return t;
}

template <class L, int ST>
__global__ void __launch_bounds__(kBlock)
k_reduce(const typename Codec<ST>::T* __restrict__ A, const typename Codec<ST>::T* __restrict__ B,
         const double* __restrict__ target, MomArgs m, int kind, double* __restrict__ partials) {
  // grid-stride over the nodes with a grid that depends on N only; each workgroup writes
  // its partial to its own slot and k_reduce_final combines the slots in a fixed order,
  // so the float64 result is the same on every run (no atomics)
  double v = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t nn = (int64_t)blockIdx.x * kBlock + threadIdx.x; nn < m.N; nn += stride) {
    const int n = (int)nn;
    if (kind == PYLBM_RED_MASS) {
#pragma unroll
      for (int i = 0; i < L::Q; ++i) {
        double fv = (double)Codec<ST>::ld(A[(size_t)i * m.N + n]);
        if (Codec<ST>::shifted) fv = fv + L::w(i);
        v = v + fv;
      }
    } else if (kind == PYLBM_RED_KINETIC) {
      double rho, u[3];
      moments_at<L, ST>(A, m, n, rho, u);
      v += 0.5 * rho * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    } else if (kind == PYLBM_RED_LOSS) {
      double rho, u[3];
      moments_at<L, ST>(A, m, n, rho, u);
#pragma unroll
      for (int a = 0; a < L::D; ++a) {
        const double e = u[a] - target[(size_t)a * m.N + n];
        v += 0.5 * e * e;
      }
    } else if (kind == PYLBM_RED_DOT) {
#pragma unroll
      for (int i = 0; i < L::Q; ++i) {
        double a = (double)Codec<ST>::ld(A[(size_t)i * m.N + n]);
        double b = (double)Codec<ST>::ld(B[(size_t)i * m.N + n]);
        if (Codec<ST>::shifted) {
          a += L::w(i);
          b += L::w(i);
        }
        v += a * b;
      }
    } else {
#pragma unroll
      for (int i = 0; i < L::Q; ++i) {
        const double a = (double)Codec<ST>::ld(A[(size_t)i * m.N + n]);
        const double b = (double)Codec<ST>::ld(B[(size_t)i * m.N + n]);
        v = max_nan(v, fabs(a - b));
      }
    }
  }
  const double t = kind == PYLBM_RED_MAXABS_DIFF ? block_reduce<true>(v) : block_reduce<false>(v);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// One workgroup: combines the n partials of k_reduce in a fixed order into *out.
template <bool MAX>
__global__ void __launch_bounds__(kBlock)
k_reduce_final(const double* __restrict__ partials, int n, double* __restrict__ out) {
  double v = 0.0;
  for (int k = threadIdx.x; k < n; k += kBlock) v = MAX ? max_nan(v, partials[k]) : v + partials[k];
  const double t = block_reduce<MAX>(v);
  if (threadIdx.x == 0) *out = t;
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Probe family of the bandwidth ceiling (pylbm_hip_stream_probe): copy with U x 16 B per
// lane, non-temporal (NT) or ordinary loads and stores; a read-only stream (the loaded
// words are folded into one value, stored only if it takes a value the probe's fill never
// gives, so the loads stay); a write-only stream.
template <int U, bool NT>
__global__ void __launch_bounds__(kBlock)
k_copy_probe(const u32x4* __restrict__ src, u32x4* __restrict__ dst, int64_t n16) {
  const int64_t base = (int64_t)blockIdx.x * (kBlock * U) + threadIdx.x;
  u32x4 v[U];
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const int64_t i = base + (int64_t)k * kBlock;
    if (i < n16) v[k] = NT ? __builtin_nontemporal_load(&src[i]) : src[i];
  }
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const int64_t i = base + (int64_t)k * kBlock;
    if (i < n16) {
      if (NT) __builtin_nontemporal_store(v[k], &dst[i]);
      else dst[i] = v[k];
    }
  }
}

template <int U>
__global__ void __launch_bounds__(kBlock)
k_read_probe(const u32x4* __restrict__ src, u32x4* __restrict__ dst, int64_t n16) {
  const int64_t base = (int64_t)blockIdx.x * (kBlock * U) + threadIdx.x;
  unsigned acc = 0;
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const int64_t i = base + (int64_t)k * kBlock;
    if (i < n16) {
      const u32x4 v = __builtin_nontemporal_load(&src[i]);
      acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
  }
  if (acc == 0x9E3779B9u) dst[blockIdx.x] = u32x4{acc, acc, acc, acc};
}

template <int U>
__global__ void __launch_bounds__(kBlock)
k_write_probe(u32x4* __restrict__ dst, int64_t n16) {
  const int64_t base = (int64_t)blockIdx.x * (kBlock * U) + threadIdx.x;
  const u32x4 v = {0x3F800000u, 0x3F800000u, 0x3F800000u, 0x3F800000u};
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const int64_t i = base + (int64_t)k * kBlock;
    if (i < n16) __builtin_nontemporal_store(v, &dst[i]);
  }
}

template <int U>
int probe_u(int kind, const void* src, void* dst, int64_t n16, hipStream_t s) {
  const int64_t per = (int64_t)kBlock * U;
  const dim3 g((unsigned)((n16 + per - 1) / per)), b(kBlock);
  switch (kind) {
    case 0: k_copy_probe<U, true><<<g, b, 0, s>>>((const u32x4*)src, (u32x4*)dst, n16); break;
    case 1: k_copy_probe<U, false><<<g, b, 0, s>>>((const u32x4*)src, (u32x4*)dst, n16); break;
    case 2: k_read_probe<U><<<g, b, 0, s>>>((const u32x4*)src, (u32x4*)dst, n16); break;
    case 3: k_write_probe<U><<<g, b, 0, s>>>((u32x4*)dst, n16); break;
    default: return PYLBM_EINVAL;
  }
  return (int)hipGetLastError();
}

// Plane-stream probe (pylbm_hip_plane_probe): the memory access pattern of a fused
// two-buffer step without its arithmetic.  Lane x reads element x of each of S planes
// (plane_elems apart, as the S populations of a node in the f[i, ...] layout), then writes
// them to the same places in dst with non-temporal stores, in the step kernel's order (all
// loads issued, then all stores).  Workgroups take the step kernel's XCD order: XCD b % 8
// serves `chunk` consecutive row segments per turn.
template <class T, int S>
__global__ void __launch_bounds__(kBlock)
k_plane_probe(const T* __restrict__ src, T* __restrict__ dst, int64_t plane_elems,
              int chunk, FastDiv div_chunk) {
  int b = (int)blockIdx.x;
  if (chunk > 0) {
    const int x = b & 7, i = b >> 3;
    const int turn = (int)div_chunk.div((uint32_t)i);
    b = (turn * 8 + x) * chunk + (i - turn * chunk);
  }
  const int64_t x = (int64_t)b * kBlock + threadIdx.x;  // plane_elems is a multiple of kBlock
  T v[S];
#pragma unroll
  for (int k = 0; k < S; ++k) v[k] = src[k * plane_elems + x];
#pragma unroll
  for (int k = 0; k < S; ++k) __builtin_nontemporal_store(v[k], &dst[k * plane_elems + x]);
}

template <class T, int S>
int plane_probe_t(const void* src, void* dst, int64_t plane_elems, hipStream_t s) {
  const int64_t nblk = plane_elems / kBlock;
  // the step kernel's order for large states: chunks of 16 (D3Q19), runs of 2 (D2Q9)
  int chunk = S == 19 ? 16 : 2;
  if (nblk % (8 * chunk) != 0) chunk = 0;
  const FastDiv dc = make_fastdiv((uint32_t)(chunk > 0 ? chunk : 1));
  k_plane_probe<T, S><<<dim3((unsigned)nblk), dim3(kBlock), 0, s>>>(
      (const T*)src, (T*)dst, plane_elems, chunk, dc);
  return (int)hipGetLastError();
}

inline unsigned nblocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

template <class L, int ST>
int aux_typed(int op, const AuxArgs& x, hipStream_t s) {
  using T = typename Codec<ST>::T;
  const int N = x.N;
  const dim3 g(nblocks(N)), b(kBlock);
  MomArgs m{};
  m.N = N;
  m.aa = x.aa;
  m.has_force = x.has_force;
  for (int a = 0; a < 3; ++a) {
    m.F[a] = (double)x.F[a];
    m.Fnz[a] = x.F[a] != 0.0f;
  }
  m.ff = x.ff;
  switch (op) {
    case kOpDecode:
      k_decode<L, ST><<<g, b, 0, s>>>((const T*)x.f, x.d_out, N, x.aa);
      break;
    case kOpEncode:
      k_encode<L, ST><<<g, b, 0, s>>>(x.d_in, (T*)x.f_out, N, x.aa);
      break;
    case kOpInit:
      k_init_eq<L, ST><<<g, b, 0, s>>>(x.d_in, x.d_in2, (T*)x.f_out, N, x.aa);
      break;
    case kOpMacro:
      k_macro<L, ST><<<g, b, 0, s>>>((const T*)x.f, m, x.d_out, x.d_out2);
      break;
    case kOpReduce: {
      const int nb = reduce_blocks(N);
      if (!x.partials) return PYLBM_EINVAL;
      k_reduce<L, ST><<<dim3(nb), b, 0, s>>>((const T*)x.f, (const T*)x.f2, x.d_in, m, x.kind,
                                              x.partials);
      if (x.kind == PYLBM_RED_MAXABS_DIFF)
        k_reduce_final<true><<<1, b, 0, s>>>(x.partials, nb, x.d_out);
      else
        k_reduce_final<false><<<1, b, 0, s>>>(x.partials, nb, x.d_out);
      break;
    }
    default:
      return PYLBM_EINVAL;
  }
  return (int)hipGetLastError();
}

template <class L>
int aux_lat(int op, const AuxArgs& x, hipStream_t s) {
  switch (x.storage) {
    case PYLBM_F32: return aux_typed<L, PYLBM_F32>(op, x, s);
    case PYLBM_F16: return aux_typed<L, PYLBM_F16>(op, x, s);
    case PYLBM_BF16: return aux_typed<L, PYLBM_BF16>(op, x, s);
    case PYLBM_F32S: return aux_typed<L, PYLBM_F32S>(op, x, s);
  }
  return PYLBM_EINVAL;
}

}  // namespace

int launch_aux(int op, const AuxArgs& x, hipStream_t s) {
  return x.lattice == PYLBM_D2Q9 ? aux_lat<D2Q9>(op, x, s) : aux_lat<D3Q19>(op, x, s);
}

int reduce_blocks(int64_t n_nodes) {
  const int64_t nb = (n_nodes + kBlock - 1) / kBlock;
  return (int)(nb < kReducePartials ? nb : kReducePartials);
}

// The allocation-free reductions keep their partials in this array, one copy per GPU for
// the life of the process.  Host code uses the address that hipGetSymbolAddress returns on
// each device (reduce_scratch), never the array's name.
static __device__ __align__(256) double scratch[kReducePartials];

namespace {
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int reduce_scratch(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= kMaxDevices) return PYLBM_ERANGE;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

int launch_plane_probe(int planes, int elem_bytes, const void* src, void* dst,
                       int64_t plane_bytes, hipStream_t s) {
  const int64_t pe = plane_bytes / elem_bytes;
  if (elem_bytes == 4) {
    if (planes == 9) return plane_probe_t<float, 9>(src, dst, pe, s);
    if (planes == 19) return plane_probe_t<float, 19>(src, dst, pe, s);
  } else if (elem_bytes == 2) {
    if (planes == 9) return plane_probe_t<uint16_t, 9>(src, dst, pe, s);
    if (planes == 19) return plane_probe_t<uint16_t, 19>(src, dst, pe, s);
  }
  return PYLBM_EINVAL;
}

int launch_probe(int kind, int unroll, const void* src, void* dst, int64_t nbytes,
                 hipStream_t s) {
  const int64_t n16 = nbytes / 16;
  switch (unroll) {
    case 1: return probe_u<1>(kind, src, dst, n16, s);
    case 2: return probe_u<2>(kind, src, dst, n16, s);
    case 4: return probe_u<4>(kind, src, dst, n16, s);
    case 8: return probe_u<8>(kind, src, dst, n16, s);
  }
  return PYLBM_EINVAL;
}

int launch_copy(const void* src, void* dst, int64_t nbytes, hipStream_t s) {
  // the fastest copy of the probe family on MI355X over 4 GiB: non-temporal, 16 B per lane
  return probe_u<1>(0, src, dst, nbytes / 16, s);
}

}  // namespace plbm
