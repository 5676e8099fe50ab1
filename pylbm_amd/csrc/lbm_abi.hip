// C ABI of libpylbm_hip.so (declared in include/pylbm_hip.h): parameter checks, kernel
// argument setup, the C-side step loop and error reporting.  Device memory is owned by the
// caller (in pylbm's hip backend: torch tensors); the only allocation here is the
// stream-ordered per-workgroup partials buffer of pylbm_hip_reduce (its allocation-free
// variant, pylbm_hip_reduce_noalloc, uses a static device array instead).
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>

#include <string>

#include "lbm_adjoint.hpp"
#include "lbm_aux.hpp"
#include "lbm_step.hpp"

namespace plbm {
int launch_step_pair_d2q9(const KArgs& a, int coll, int storage, int fmode, const void* src,
                          void* dst, hipStream_t s);
int launch_sweep_probe_d2q9(const KArgs& a, int elem_bytes, const void* src, void* dst,
                            hipStream_t s);
template <class L>
int launch_step(const KArgs& a, int coll, int storage, int pattern, int fmode,
                const void* src, void* dst, hipStream_t s);
extern template int launch_step<D2Q9>(const KArgs&, int, int, int, int, const void*, void*,
                                      hipStream_t);
extern template int launch_step<D3Q19>(const KArgs&, int, int, int, int, const void*, void*,
                                       hipStream_t);
template <class L>
int launch_step_geo(const KArgs& a, int coll, int storage, int fmode, const void* src,
                    void* dst, hipStream_t s);
template <class L>
int launch_links(const uint8_t* solid, void* links, int nx, int ny, int nz, hipStream_t s);
template <class L>
int launch_link_groups(const void* links, uint32_t* groups, int nx, int rows, int perx,
                       hipStream_t s);
extern template int launch_step_geo<D2Q9>(const KArgs&, int, int, int, const void*, void*,
                                          hipStream_t);
extern template int launch_step_geo<D3Q19>(const KArgs&, int, int, int, const void*, void*,
                                           hipStream_t);
extern template int launch_links<D2Q9>(const uint8_t*, void*, int, int, int, hipStream_t);
extern template int launch_links<D3Q19>(const uint8_t*, void*, int, int, int, hipStream_t);
extern template int launch_link_groups<D2Q9>(const void*, uint32_t*, int, int, int,
                                             hipStream_t);
extern template int launch_link_groups<D3Q19>(const void*, uint32_t*, int, int, int,
                                              hipStream_t);
}  // namespace plbm

using namespace plbm;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(int code, const char* what) {
  if (code == 0) return 0;
  if (code >= PYLBM_EINVAL && code <= PYLBM_ENOTSUP)
    return fail(code, "%s: unsupported parameter combination", what);
  return fail(code, "%s: HIP error %d (%s)", what, code,
              hipGetErrorString((hipError_t)code));
}

int check_params(const pylbm_params* p, const char* what) {
  if (!p) return fail(PYLBM_EINVAL, "%s: params is NULL", what);
  if (p->lattice != PYLBM_D2Q9 && p->lattice != PYLBM_D3Q19)
    return fail(PYLBM_EINVAL, "%s: unknown lattice %d", what, p->lattice);
  if (p->collision < PYLBM_BGK || p->collision > PYLBM_MRT)
    return fail(PYLBM_EINVAL, "%s: unknown collision %d", what, p->collision);
  if (p->collision == PYLBM_MRT && p->lattice != PYLBM_D2Q9)
    return fail(PYLBM_ENOTSUP, "%s: MRT is implemented for D2Q9 only", what);
  if (p->storage < PYLBM_F32 || p->storage > PYLBM_F32S)
    return fail(PYLBM_EINVAL, "%s: unknown storage %d", what, p->storage);
  if (p->aa != 0 && p->aa != 1) return fail(PYLBM_EINVAL, "%s: aa must be 0 or 1", what);
  if (p->nx < 1 || p->ny < 1 || p->nz < 1)
    return fail(PYLBM_EINVAL, "%s: sizes must be >= 1 (got %d x %d x %d)", what, p->nx,
                p->ny, p->nz);
  if (p->lattice == PYLBM_D2Q9 && p->nz != 1)
    return fail(PYLBM_EINVAL, "%s: D2Q9 needs nz == 1", what);
  const int64_t N = (int64_t)p->nx * p->ny * p->nz;
  if (N != p->n_nodes)
    return fail(PYLBM_EINVAL, "%s: n_nodes %lld != nx*ny*nz = %lld", what,
                (long long)p->n_nodes, (long long)N);
  if (N >= ((int64_t)1 << 31) - kBlock)
    return fail(PYLBM_ERANGE, "%s: %lld nodes exceed the 32-bit in-plane index", what,
                (long long)N);
  const int64_t item = (p->storage == PYLBM_F16 || p->storage == PYLBM_BF16) ? 2 : 4;
  if (N * item >= ((int64_t)1 << 32))
    return fail(PYLBM_ERANGE, "%s: a %lld-node plane exceeds the 4 GiB byte offsets of the "
                "kernels", what, (long long)N);
  const int last = p->lattice == PYLBM_D2Q9 ? 1 : 2;
  if (p->ulid != 0.0f && p->periodic[last])
    return fail(PYLBM_EINVAL, "%s: a moving lid needs a wall on the last axis", what);
  // fp32 side arrays (the per-cell force field) use 32-bit byte offsets as well
  if (p->force_field && N * 4 >= ((int64_t)1 << 32))
    return fail(PYLBM_ERANGE, "%s: a %lld-node fp32 force-field plane exceeds the 4 GiB "
                "byte offsets of the kernels", what, (long long)N);
  if (p->lattice == PYLBM_D2Q9 && p->force_field == nullptr && p->force[2] != 0.0f)
    return fail(PYLBM_EINVAL, "%s: D2Q9 has no z force", what);
  if (p->lam_storage != PYLBM_F32 && p->lam_storage != PYLBM_F16)
    return fail(PYLBM_EINVAL, "%s: lam_storage must be PYLBM_F32 or PYLBM_F16", what);
  if (p->outlet != 0 && p->outlet != 1) return fail(PYLBM_EINVAL, "%s: outlet must be 0 or 1", what);
  if (p->link_groups && !p->links)
    return fail(PYLBM_EINVAL, "%s: link_groups needs links", what);
  if (p->links || p->inlet || p->outlet) {
    if ((p->inlet || p->outlet) && p->lattice != PYLBM_D2Q9)
      return fail(PYLBM_ENOTSUP, "%s: the inlet and the outlet are D2Q9 only", what);
    if (p->outlet && p->nx < 2) return fail(PYLBM_EINVAL, "%s: the outlet needs nx >= 2", what);
    if (p->aa)
      return fail(PYLBM_ENOTSUP, "%s: obstacles, inlet and outlet run on the two-buffer "
                  "layout only", what);
    const int64_t q = p->lattice == PYLBM_D2Q9 ? 9 : 19;
    if (q * N * item >= ((int64_t)1 << 32))
      return fail(PYLBM_ERANGE, "%s: with obstacles, inlet or outlet the population array "
                  "must span less than 4 GiB", what);
  }
  return 0;
}

bool has_geometry(const pylbm_params* p) { return p->links || p->inlet || p->outlet; }

int make_kargs(const pylbm_params* p, KArgs& a, const char* what) {
  a.nx = p->nx;
  a.ny = p->ny;
  a.nz = p->nz;
  a.N = (int)p->n_nodes;
  const int64_t rows = (int64_t)p->ny * p->nz;
  a.rows = (int)rows;
  int64_t nblk;
  if (p->nx <= 128) {
    // short rows: 2 rows of 128 lanes or 4 rows of 64 lanes per workgroup
    a.sw_log2 = p->nx <= 64 ? 6 : 7;
    a.rpw_log2 = 8 - a.sw_log2;
    a.nseg = 1;
    nblk = (rows + (1 << a.rpw_log2) - 1) >> a.rpw_log2;
  } else {
    a.sw_log2 = 8;
    a.rpw_log2 = 0;
    a.nseg = (p->nx + kBlock - 1) / kBlock;
    nblk = (int64_t)a.nseg * rows;
  }
  if (nblk >= ((int64_t)1 << 31)) return fail(PYLBM_ERANGE, "%s: grid too large", what);
  a.nblk = (int)nblk;
  const int dims = p->lattice == PYLBM_D2Q9 ? 2 : 3;
  a.per = 0;
  for (int k = 0; k < dims; ++k)
    if (p->periodic[k]) a.per |= 1 << k;
  // workgroup order over the XCDs (remap_block, set_xcd_order)
  const int64_t q = p->lattice == PYLBM_D2Q9 ? 9 : 19;
  const int64_t item = (p->storage == PYLBM_F16 || p->storage == PYLBM_BF16) ? 2 : 4;
  const bool cached = 2 * q * (int64_t)p->n_nodes * item <= ((int64_t)256 << 20);
  set_xcd_order(a, cached ? kOrderSlabs
                          : (p->lattice == PYLBM_D3Q19 ? kOrderChunk16 : kOrderRun));
  a.reverse = 0;
  a.resident = cached ? 1 : 0;
  a.div_nseg = make_fastdiv((uint32_t)a.nseg);
  a.div_ny = make_fastdiv((uint32_t)p->ny);
  a.ulid = p->ulid;
  a.fx = p->force[0];
  a.fy = p->force[1];
  a.fz = p->lattice == PYLBM_D2Q9 ? 0.0f : p->force[2];
  a.r0 = p->rates[0];
  a.r1 = p->rates[1];
  a.r2 = p->rates[2];
  a.r3 = p->rates[3];
  a.ff = p->force_field;
  a.links = p->links;
  a.inlet = p->inlet;
  a.outlet = p->outlet;
  a.rec = nullptr;
  a.lgroups = p->link_groups;
  a.gpr = (p->nx + 63) / 64;
  return 0;
}

// the two-steps-per-launch pattern: D2Q9 box problems (periodic axes or bounce-back walls,
// the lid; no obstacles, inlet or outlet), the state under 4 GiB (the kernel addresses it
// through one buffer descriptor)
bool pair_steps_apply(const pylbm_params* p) {
  const uint64_t ts = (p->storage == PYLBM_F16 || p->storage == PYLBM_BF16) ? 2 : 4;
  return p->lattice == PYLBM_D2Q9 && !p->links && !p->inlet && !p->outlet && !p->aa &&
         9ull * (uint64_t)p->n_nodes * ts < (1ull << 32);
}

int step_impl(const pylbm_params* p, const KArgs& a, const void* src, void* dst, int pattern,
              hipStream_t s) {
  const int fmode = p->force_field != nullptr ? kForceField
                    : (a.fx != 0.0f || a.fy != 0.0f || a.fz != 0.0f) ? kForceUniform
                                                                     : kForceNone;
  if (pattern == PYLBM_AB2) {
    if (!pair_steps_apply(p)) return PYLBM_ENOTSUP;
    return launch_step_pair_d2q9(a, p->collision, p->storage, fmode, src, dst, s);
  }
  if (has_geometry(p)) {
    if (pattern != PYLBM_AB) return PYLBM_ENOTSUP;
    if (p->lattice == PYLBM_D2Q9)
      return launch_step_geo<D2Q9>(a, p->collision, p->storage, fmode, src, dst, s);
    return launch_step_geo<D3Q19>(a, p->collision, p->storage, fmode, src, dst, s);
  }
  if (p->lattice == PYLBM_D2Q9)
    return launch_step<D2Q9>(a, p->collision, p->storage, pattern, fmode, src, dst, s);
  return launch_step<D3Q19>(a, p->collision, p->storage, pattern, fmode, src, dst, s);
}

// The record mode of the forward steps and the adjoint: the record array is N x 16 bytes and
// the kernels address it, the adjoint states and every other array through one descriptor
// each, so a q-plane fp32 array must span less than 4 GiB.
int check_records(const pylbm_params* p, const char* what) {
  const int64_t q = p->lattice == PYLBM_D2Q9 ? 9 : 19;
  if (q * p->n_nodes * 4 >= ((int64_t)1 << 32))
    return fail(PYLBM_ERANGE, "%s: moment records need q * N * 4 bytes under 4 GiB", what);
  if (p->collision == PYLBM_MRT)
    return fail(PYLBM_ENOTSUP, "%s: moment records cover BGK and TRT", what);
  if (has_geometry(p))
    return fail(PYLBM_ENOTSUP, "%s: moment records cover box geometries only", what);
  return 0;
}

// The stepping loop of pylbm_hip_run_store / pylbm_hip_run_record: step k reads the
// previous state (src for k = 0) and writes dst[k]; with rec, a non-NULL rec[k] also
// receives step k's moment records.
int run_store_impl(const pylbm_params* p, const void* src, void* const* dst, void* const* rec,
                   int64_t nsteps, void* stream, const char* what) {
  if (int e = check_params(p, what)) return e;
  if (p->aa) return fail(PYLBM_EINVAL, "%s: two-buffer layout only", what);
  if (nsteps < 0) return fail(PYLBM_EINVAL, "%s: nsteps < 0", what);
  if (nsteps > 0 && (!src || !dst)) return fail(PYLBM_EINVAL, "%s: NULL buffer", what);
  if (rec)
    if (int e = check_records(p, what)) return e;
  KArgs a;
  if (int e = make_kargs(p, a, what)) return e;
  KArgs ar = a;
  ar.reverse = 1;
  const void* cur = src;
  for (int64_t k = 0; k < nsteps; ++k) {
    if (!dst[k] || dst[k] == cur)
      return fail(PYLBM_EINVAL, "%s: dst[%lld] is NULL or equals its source", what,
                  (long long)k);
    KArgs ak = ((k & 1) && a.resident) ? ar : a;
    ak.rec = rec ? static_cast<float*>(rec[k]) : nullptr;
    if (int e = step_impl(p, ak, cur, dst[k], PYLBM_AB, (hipStream_t)stream))
      return hip_fail(e, what);
    cur = dst[k];
  }
  return 0;
}

AuxArgs aux_from(const pylbm_params* p) {
  AuxArgs x{};
  x.lattice = p->lattice;
  x.storage = p->storage;
  x.aa = p->aa;
  x.N = (int)p->n_nodes;
  x.ff = p->force_field;
  x.has_force = (p->force_field != nullptr || p->force[0] != 0.0f || p->force[1] != 0.0f ||
                 p->force[2] != 0.0f) ? 1 : 0;
  for (int a = 0; a < 3; ++a) x.F[a] = p->force[a];
  return x;
}

}  // namespace

extern "C" {

int pylbm_hip_abi_version(void) { return PYLBM_HIP_ABI_VERSION; }

const char* pylbm_hip_last_error(void) { return g_err.c_str(); }

int pylbm_hip_step(const pylbm_params* p, const void* src, void* dst, int pattern,
                   void* stream) {
  if (int e = check_params(p, "pylbm_hip_step")) return e;
  if (!src || !dst) return fail(PYLBM_EINVAL, "pylbm_hip_step: NULL buffer");
  if (pattern == PYLBM_AB || pattern == PYLBM_AB2) {
    if (src == dst) return fail(PYLBM_EINVAL, "pylbm_hip_step: AB needs src != dst");
    if (p->aa) return fail(PYLBM_EINVAL, "pylbm_hip_step: AB pattern on an AA-layout state");
    if (pattern == PYLBM_AB2 && !pair_steps_apply(p))
      return fail(PYLBM_ENOTSUP, "pylbm_hip_step: two steps per launch run on D2Q9 box "
                  "problems with a state under 4 GiB only");
  } else if (pattern == PYLBM_AA_ODD || pattern == PYLBM_AA_EVEN) {
    if (src != dst) return fail(PYLBM_EINVAL, "pylbm_hip_step: AA patterns work in place");
  } else {
    return fail(PYLBM_EINVAL, "pylbm_hip_step: unknown pattern %d", pattern);
  }
  KArgs a;
  if (int e = make_kargs(p, a, "pylbm_hip_step")) return e;
  return hip_fail(step_impl(p, a, src, dst, pattern, (hipStream_t)stream), "pylbm_hip_step");
}

int pylbm_hip_run(const pylbm_params* p, void* a_buf, void* b_buf, int64_t nsteps,
                  void* stream) {
  if (int e = check_params(p, "pylbm_hip_run")) return e;
  if (nsteps < 0) return fail(PYLBM_EINVAL, "pylbm_hip_run: nsteps < 0");
  if (!a_buf) return fail(PYLBM_EINVAL, "pylbm_hip_run: NULL buffer");
  KArgs a;
  if (int e = make_kargs(p, a, "pylbm_hip_run")) return e;
  KArgs ar = a;
  ar.reverse = 1;
  hipStream_t s = (hipStream_t)stream;
  if (p->aa) {
    if (nsteps % 2) return fail(PYLBM_EINVAL, "pylbm_hip_run: AA advances in pairs of steps");
    // AA: the sweep direction alternates every launch (remap_block)
    for (int64_t k = 0; k < nsteps; k += 2) {
      if (int e = step_impl(p, a, a_buf, a_buf, PYLBM_AA_ODD, s))
        return hip_fail(e, "pylbm_hip_run (aa-odd)");
      if (int e = step_impl(p, ar, a_buf, a_buf, PYLBM_AA_EVEN, s))
        return hip_fail(e, "pylbm_hip_run (aa-even)");
    }
    return 0;
  }
  if (!b_buf || b_buf == a_buf)
    return fail(PYLBM_EINVAL, "pylbm_hip_run: two-buffer stepping needs a distinct b");
  void* src = a_buf;
  void* dst = b_buf;
  // two buffers: the direction alternates for resident states and for the geometry
  // kernel (it re-reads the same link masks every step: 4096^2 channel fp16 / bf16 +3.6%,
  // fp32 +0.8%); large box states gain nothing, their stores stream past the Infinity
  // Cache (alternating measured 1-2.5% slower, with temporal stores 3-4% slower)
  const bool alternate = a.resident || has_geometry(p);
  for (int64_t k = 0; k < nsteps; ++k) {
    if (int e = step_impl(p, ((k & 1) && alternate) ? ar : a, src, dst, PYLBM_AB, s))
      return hip_fail(e, "pylbm_hip_run");
    void* t = src;
    src = dst;
    dst = t;
  }
  return 0;
}

int pylbm_hip_run_pairs(const pylbm_params* p, void* a_buf, void* b_buf, int64_t nsteps,
                        void* stream) {
  if (int e = check_params(p, "pylbm_hip_run_pairs")) return e;
  if (nsteps < 0 || nsteps % 2)
    return fail(PYLBM_EINVAL, "pylbm_hip_run_pairs: nsteps must be even and >= 0");
  if (!a_buf || !b_buf || a_buf == b_buf)
    return fail(PYLBM_EINVAL, "pylbm_hip_run_pairs: needs two distinct buffers");
  if (!pair_steps_apply(p))
    return fail(PYLBM_ENOTSUP, "pylbm_hip_run_pairs: D2Q9 box problems with a state under "
                "4 GiB only");
  KArgs a;
  if (int e = make_kargs(p, a, "pylbm_hip_run_pairs")) return e;
  void* src = a_buf;
  void* dst = b_buf;
  for (int64_t k = 0; k < nsteps; k += 2) {
    if (int e = step_impl(p, a, src, dst, PYLBM_AB2, (hipStream_t)stream))
      return hip_fail(e, "pylbm_hip_run_pairs");
    void* t = src;
    src = dst;
    dst = t;
  }
  return 0;
}

int pylbm_hip_run_store(const pylbm_params* p, const void* src, void* const* dst,
                        int64_t nsteps, void* stream) {
  return run_store_impl(p, src, dst, nullptr, nsteps, stream, "pylbm_hip_run_store");
}

int pylbm_hip_run_record(const pylbm_params* p, const void* src, void* const* dst,
                         void* const* rec, int64_t nsteps, void* stream) {
  return run_store_impl(p, src, dst, rec, nsteps, stream, "pylbm_hip_run_record");
}

int pylbm_hip_encode(const pylbm_params* p, const double* canon, void* f, void* stream) {
  if (int e = check_params(p, "pylbm_hip_encode")) return e;
  if (!canon || !f) return fail(PYLBM_EINVAL, "pylbm_hip_encode: NULL buffer");
  AuxArgs x = aux_from(p);
  x.d_in = canon;
  x.f_out = f;
  return hip_fail(launch_aux(kOpEncode, x, (hipStream_t)stream), "pylbm_hip_encode");
}

int pylbm_hip_decode(const pylbm_params* p, const void* f, double* canon, void* stream) {
  if (int e = check_params(p, "pylbm_hip_decode")) return e;
  if (!canon || !f) return fail(PYLBM_EINVAL, "pylbm_hip_decode: NULL buffer");
  AuxArgs x = aux_from(p);
  x.f = f;
  x.d_out = canon;
  return hip_fail(launch_aux(kOpDecode, x, (hipStream_t)stream), "pylbm_hip_decode");
}

int pylbm_hip_init_equilibrium(const pylbm_params* p, const double* rho, const double* u,
                               void* f, void* stream) {
  if (int e = check_params(p, "pylbm_hip_init_equilibrium")) return e;
  if (!f) return fail(PYLBM_EINVAL, "pylbm_hip_init_equilibrium: NULL buffer");
  AuxArgs x = aux_from(p);
  x.d_in = rho;
  x.d_in2 = u;
  x.f_out = f;
  return hip_fail(launch_aux(kOpInit, x, (hipStream_t)stream), "pylbm_hip_init_equilibrium");
}

int pylbm_hip_macroscopic(const pylbm_params* p, const void* f, double* rho, double* u,
                          void* stream) {
  if (int e = check_params(p, "pylbm_hip_macroscopic")) return e;
  if (!f) return fail(PYLBM_EINVAL, "pylbm_hip_macroscopic: NULL buffer");
  if (!rho && !u) return 0;
  AuxArgs x = aux_from(p);
  x.f = f;
  x.d_out = rho;
  x.d_out2 = u;
  return hip_fail(launch_aux(kOpMacro, x, (hipStream_t)stream), "pylbm_hip_macroscopic");
}

static int check_reduce(const pylbm_params* p, int kind, const void* f, const void* other,
                        const double* out, const char* what) {
  if (int e = check_params(p, what)) return e;
  if (!f || !out) return fail(PYLBM_EINVAL, "%s: NULL buffer", what);
  if (kind < PYLBM_RED_MASS || kind > PYLBM_RED_DOT)
    return fail(PYLBM_EINVAL, "%s: unknown kind %d", what, kind);
  if ((kind == PYLBM_RED_MAXABS_DIFF || kind == PYLBM_RED_DOT) && !other)
    return fail(PYLBM_EINVAL, "%s: this kind needs a second state", what);
  if (kind == PYLBM_RED_LOSS && !other)
    return fail(PYLBM_EINVAL, "%s: LOSS needs the target velocity", what);
  return 0;
}

static AuxArgs reduce_args(const pylbm_params* p, int kind, const void* f, const void* other,
                           double* out) {
  AuxArgs x = aux_from(p);
  x.f = f;
  x.f2 = kind == PYLBM_RED_LOSS ? nullptr : other;
  x.d_in = kind == PYLBM_RED_LOSS ? static_cast<const double*>(other) : nullptr;
  x.kind = kind;
  x.d_out = out;
  return x;
}

int pylbm_hip_reduce(const pylbm_params* p, int kind, const void* f, const void* other,
                     double* out, void* stream) {
  if (int e = check_reduce(p, kind, f, other, out, "pylbm_hip_reduce")) return e;
  AuxArgs x = reduce_args(p, kind, f, other, out);
  // the per-workgroup partials live for the two launches only: a stream-ordered
  // allocation, so concurrent reductions on other streams never share them
  hipStream_t s = (hipStream_t)stream;
  void* partials = nullptr;
  if (hipError_t e = hipMallocAsync(&partials, sizeof(double) * reduce_blocks(p->n_nodes), s))
    return hip_fail((int)e, "pylbm_hip_reduce (partials)");
  x.partials = static_cast<double*>(partials);
  const int e = launch_aux(kOpReduce, x, s);
  const hipError_t ef = hipFreeAsync(partials, s);
  if (e) return hip_fail(e, "pylbm_hip_reduce");
  return hip_fail((int)ef, "pylbm_hip_reduce (partials)");
}

int pylbm_hip_reduce_noalloc(const pylbm_params* p, int kind, const void* f,
                             const void* other, double* out, void* stream) {
  if (int e = check_reduce(p, kind, f, other, out, "pylbm_hip_reduce_noalloc")) return e;
  // the grid is capped at kReducePartials workgroups, so every size check_params accepts
  // fits the static array; the check guards that invariant
  const int nb = reduce_blocks(p->n_nodes);
  if (nb > kReducePartials)
    return fail(PYLBM_ERANGE, "pylbm_hip_reduce_noalloc: %d partials exceed the %d of the "
                "static scratch", nb, kReducePartials);
  double* scratch_ptr = nullptr;
  if (int e = reduce_scratch(&scratch_ptr))
    return hip_fail(e, "pylbm_hip_reduce_noalloc (scratch address)");
  AuxArgs x = reduce_args(p, kind, f, other, out);
  x.partials = scratch_ptr;
  return hip_fail(launch_aux(kOpReduce, x, (hipStream_t)stream), "pylbm_hip_reduce_noalloc");
}

int pylbm_hip_links(const pylbm_params* p, const uint8_t* solid, void* links, void* stream) {
  if (int e = check_params(p, "pylbm_hip_links")) return e;
  if (!solid || !links) return fail(PYLBM_EINVAL, "pylbm_hip_links: NULL buffer");
  if (p->n_nodes == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  const int e = p->lattice == PYLBM_D2Q9 ? launch_links<D2Q9>(solid, links, p->nx, p->ny, p->nz, s)
                                         : launch_links<D3Q19>(solid, links, p->nx, p->ny, p->nz, s);
  return hip_fail(e, "pylbm_hip_links");
}

int pylbm_hip_link_groups(const pylbm_params* p, const void* links, uint32_t* groups,
                          void* stream) {
  if (int e = check_params(p, "pylbm_hip_link_groups")) return e;
  if (!links || !groups) return fail(PYLBM_EINVAL, "pylbm_hip_link_groups: NULL buffer");
  if (p->n_nodes == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  const int rows = p->ny * p->nz;
  const int perx = p->periodic[0] ? 1 : 0;
  const int e = p->lattice == PYLBM_D2Q9
                    ? launch_link_groups<D2Q9>(links, groups, p->nx, rows, perx, s)
                    : launch_link_groups<D3Q19>(links, groups, p->nx, rows, perx, s);
  return hip_fail(e, "pylbm_hip_link_groups");
}

int pylbm_hip_copy(const void* src, void* dst, int64_t nbytes, void* stream) {
  if (!src || !dst) return fail(PYLBM_EINVAL, "pylbm_hip_copy: NULL buffer");
  if (nbytes < 0 || nbytes % 16)
    return fail(PYLBM_EINVAL, "pylbm_hip_copy: nbytes must be a non-negative multiple of 16");
  if (nbytes == 0) return 0;
  return hip_fail(launch_copy(src, dst, nbytes, (hipStream_t)stream), "pylbm_hip_copy");
}

int pylbm_hip_stream_probe(int kind, int unroll, const void* src, void* dst, int64_t nbytes,
                           void* stream) {
  if (kind < 0 || kind > 3) return fail(PYLBM_EINVAL, "pylbm_hip_stream_probe: kind 0..3");
  if (unroll != 1 && unroll != 2 && unroll != 4 && unroll != 8)
    return fail(PYLBM_EINVAL, "pylbm_hip_stream_probe: unroll must be 1, 2, 4 or 8");
  if ((kind != 3 && !src) || !dst) return fail(PYLBM_EINVAL, "pylbm_hip_stream_probe: NULL buffer");
  if (nbytes < 0 || nbytes % 16)
    return fail(PYLBM_EINVAL, "pylbm_hip_stream_probe: nbytes must be a non-negative multiple "
                "of 16");
  if (kind == 2 && nbytes / 16 > (int64_t)kBlock * unroll * ((int64_t)1 << 31))
    return fail(PYLBM_ERANGE, "pylbm_hip_stream_probe: too many workgroups");
  if (nbytes == 0) return 0;
  return hip_fail(launch_probe(kind, unroll, src, dst, nbytes, (hipStream_t)stream),
                  "pylbm_hip_stream_probe");
}

int pylbm_hip_plane_probe(int planes, int elem_bytes, const void* src, void* dst,
                          int64_t plane_bytes, void* stream) {
  if (planes != 9 && planes != 19)
    return fail(PYLBM_EINVAL, "pylbm_hip_plane_probe: planes must be 9 or 19");
  if (elem_bytes != 2 && elem_bytes != 4)
    return fail(PYLBM_EINVAL, "pylbm_hip_plane_probe: elem_bytes must be 2 or 4");
  if (!src || !dst) return fail(PYLBM_EINVAL, "pylbm_hip_plane_probe: NULL buffer");
  if (plane_bytes < 0 || plane_bytes % ((int64_t)kBlock * elem_bytes))
    return fail(PYLBM_EINVAL, "pylbm_hip_plane_probe: plane_bytes must be a non-negative "
                "multiple of 256 elements");
  if (plane_bytes / elem_bytes / kBlock >= ((int64_t)1 << 31))
    return fail(PYLBM_ERANGE, "pylbm_hip_plane_probe: too many workgroups");
  if (plane_bytes == 0) return 0;
  return hip_fail(plbm::launch_plane_probe(planes, elem_bytes, src, dst, plane_bytes,
                                           (hipStream_t)stream),
                  "pylbm_hip_plane_probe");
}

int pylbm_hip_sweep_probe(int elem_bytes, const void* src, void* dst, int nx, int ny,
                          void* stream) {
  if (elem_bytes != 2 && elem_bytes != 4)
    return fail(PYLBM_EINVAL, "pylbm_hip_sweep_probe: elem_bytes must be 2 or 4");
  if (!src || !dst || src == dst)
    return fail(PYLBM_EINVAL, "pylbm_hip_sweep_probe: needs two distinct buffers");
  if (nx < 1 || ny < 1) return fail(PYLBM_EINVAL, "pylbm_hip_sweep_probe: nx, ny must be >= 1");
  if (9 * (int64_t)nx * ny * elem_bytes >= ((int64_t)1 << 32))
    return fail(PYLBM_ERANGE, "pylbm_hip_sweep_probe: the state must span less than 4 GiB");
  plbm::KArgs a{};
  a.nx = nx;
  a.ny = ny;
  a.nz = 1;
  a.N = nx * ny;
  return hip_fail(plbm::launch_sweep_probe_d2q9(a, elem_bytes, src, dst, (hipStream_t)stream),
                  "pylbm_hip_sweep_probe");
}

int64_t pylbm_hip_adjoint_workspace(const pylbm_params* p) {
  if (check_params(p, "pylbm_hip_adjoint_workspace")) return -1;
  KArgs a;
  if (make_kargs(p, a, "pylbm_hip_adjoint_workspace")) return -1;
  return adjoint_partials(a);
}

static int check_adjoint(const pylbm_params* p, const char* what) {
  if (int e = check_params(p, what)) return e;
  // lam (fp32 by default), dF and grad_u are fp32 planes indexed with 32-bit byte offsets
  if (p->n_nodes * 4 >= ((int64_t)1 << 32))
    return fail(PYLBM_ERANGE, "%s: a %lld-node fp32 adjoint plane exceeds the 4 GiB byte "
                "offsets of the kernels", what, (long long)p->n_nodes);
  if (p->aa) return fail(PYLBM_EINVAL, "%s: the adjoint runs on the two-buffer layout", what);
  if (p->collision == PYLBM_MRT)
    return fail(PYLBM_ENOTSUP, "%s: the adjoint covers BGK and TRT", what);
  if (has_geometry(p))
    return fail(PYLBM_ENOTSUP, "%s: the adjoint covers box geometries only", what);
  return 0;
}

static int adjoint_step_impl(const pylbm_params* p, const void* f_t, bool rec,
                             const void* lam_in, void* lam_out, float* dF, double* dtau_part,
                             void* stream, const char* what) {
  if (int e = check_adjoint(p, what)) return e;
  if (rec)
    if (int e = check_records(p, what)) return e;
  if (!f_t || !lam_in || !lam_out) return fail(PYLBM_EINVAL, "%s: NULL buffer", what);
  if (lam_in == lam_out) return fail(PYLBM_EINVAL, "%s: lam_in and lam_out must differ", what);
  if (p->force_field && !dF) return fail(PYLBM_EINVAL, "%s: a force field needs dF", what);
  KArgs a;
  if (int e = make_kargs(p, a, what)) return e;
  return hip_fail(launch_adjoint_step(p->lattice, a, p->collision, p->storage, p->lam_storage,
                                      PYLBM_AB, p->force_field != nullptr, rec, f_t, lam_in,
                                      lam_out, dF, dtau_part, (hipStream_t)stream),
                  what);
}

static int adjoint_run_impl(const pylbm_params* p, const void* const* inputs, bool rec,
                            int64_t nsteps, void* lam_a, void* lam_b, float* dF,
                            double* dtau_part, int* which, void* stream, const char* what) {
  if (int e = check_adjoint(p, what)) return e;
  if (rec)
    if (int e = check_records(p, what)) return e;
  if (nsteps < 0 || !which) return fail(PYLBM_EINVAL, "%s: bad arguments", what);
  if (nsteps > 0 && (!inputs || !lam_a || !lam_b || lam_a == lam_b))
    return fail(PYLBM_EINVAL, "%s: NULL or aliased buffer", what);
  if (p->force_field && !dF) return fail(PYLBM_EINVAL, "%s: a force field needs dF", what);
  KArgs a;
  if (int e = make_kargs(p, a, what)) return e;
  KArgs ar = a;
  ar.reverse = 1;
  for (int64_t k = 0; k < nsteps; ++k)
    if (!inputs[k]) return fail(PYLBM_EINVAL, "%s: input %lld is NULL", what, (long long)k);
  void* lin = lam_a;
  void* lout = lam_b;
  const bool field = p->force_field != nullptr;
  if (rec && nsteps >= 2 &&
      adjoint_pair_applies(a, p->lattice, p->collision, p->lam_storage, field)) {
    // two steps per launch (lam_a -> lam_b -> lam_a ...), an odd last step two-buffer
    int64_t k = nsteps - 1;
    for (; k >= 1; k -= 2) {
      if (int e = launch_adjoint_pair(a, p->collision, inputs[k], inputs[k - 1], lin, lout,
                                      (hipStream_t)stream))
        return hip_fail(e, what);
      void* t = lin;
      lin = lout;
      lout = t;
    }
    if (k == 0) {
      if (int e = launch_adjoint_step(p->lattice, a, p->collision, p->storage, p->lam_storage,
                                      PYLBM_AB, field, true, inputs[0], lin, lout, dF,
                                      dtau_part, (hipStream_t)stream))
        return hip_fail(e, what);
      void* t = lin;
      lin = lout;
      lout = t;
    }
    *which = (lin == lam_a) ? 0 : 1;
    return 0;
  }
  // The steps run in pairs in place in lam_a (the AA pattern of the adjoint, k_adjoint):
  // one adjoint-state array instead of two, which a state resident in the Infinity Cache
  // keeps there.  An odd last step goes lam_a -> lam_b.
  for (int64_t k = nsteps - 1; k >= 0; --k) {
    // the sweep direction alternates every step for resident states (remap_block)
    const bool rev = ((nsteps - 1 - k) & 1) && a.resident;
    const bool pair = (nsteps - 1 - k) < nsteps - nsteps % 2;
    const int pat = !pair ? PYLBM_AB : (((nsteps - 1 - k) & 1) ? PYLBM_AA_ODD : PYLBM_AA_EVEN);
    if (int e = launch_adjoint_step(p->lattice, rev ? ar : a, p->collision, p->storage,
                                    p->lam_storage, pat, p->force_field != nullptr, rec,
                                    inputs[k], lin, pair ? lin : lout, dF, dtau_part,
                                    (hipStream_t)stream))
      return hip_fail(e, what);
    if (!pair) {
      void* t = lin;
      lin = lout;
      lout = t;
    }
  }
  *which = (lin == lam_a) ? 0 : 1;
  return 0;
}

int pylbm_hip_adjoint_seed(const pylbm_params* p, const void* f_T, const float* grad_u,
                           void* lam, float* dF, void* stream) {
  if (int e = check_adjoint(p, "pylbm_hip_adjoint_seed")) return e;
  if (!f_T || !grad_u || !lam) return fail(PYLBM_EINVAL, "pylbm_hip_adjoint_seed: NULL buffer");
  if (p->force_field && !dF)
    return fail(PYLBM_EINVAL, "pylbm_hip_adjoint_seed: a force field needs dF");
  KArgs a;
  if (int e = make_kargs(p, a, "pylbm_hip_adjoint_seed")) return e;
  return hip_fail(launch_adjoint_seed(p->lattice, a, p->storage, p->lam_storage,
                                      p->force_field != nullptr, f_T, grad_u, lam, dF,
                                      (hipStream_t)stream),
                  "pylbm_hip_adjoint_seed");
}

int pylbm_hip_adjoint_step(const pylbm_params* p, const void* f_t, const void* lam_in,
                           void* lam_out, float* dF, double* dtau_part, void* stream) {
  return adjoint_step_impl(p, f_t, false, lam_in, lam_out, dF, dtau_part, stream,
                           "pylbm_hip_adjoint_step");
}

int pylbm_hip_adjoint_step_record(const pylbm_params* p, const void* rec_t, const void* lam_in,
                                  void* lam_out, float* dF, void* stream) {
  return adjoint_step_impl(p, rec_t, true, lam_in, lam_out, dF, nullptr, stream,
                           "pylbm_hip_adjoint_step_record");
}

int pylbm_hip_adjoint_run(const pylbm_params* p, const void* const* states, int64_t nsteps,
                          void* lam_a, void* lam_b, float* dF, double* dtau_part, int* which,
                          void* stream) {
  return adjoint_run_impl(p, states, false, nsteps, lam_a, lam_b, dF, dtau_part, which, stream,
                          "pylbm_hip_adjoint_run");
}

int pylbm_hip_adjoint_run_record(const pylbm_params* p, const void* const* recs, int64_t nsteps,
                                 void* lam_a, void* lam_b, float* dF, int* which, void* stream) {
  return adjoint_run_impl(p, recs, true, nsteps, lam_a, lam_b, dF, nullptr, which, stream,
                          "pylbm_hip_adjoint_run_record");
}

int pylbm_hip_adjoint_init(const pylbm_params* p, const double* rho0, const double* u0,
                           const void* lam0, double* du0, double* drho0, void* stream) {
  if (int e = check_params(p, "pylbm_hip_adjoint_init")) return e;
  if (!lam0) return fail(PYLBM_EINVAL, "pylbm_hip_adjoint_init: NULL lam0");
  if (!du0 && !drho0) return 0;
  return hip_fail(launch_adjoint_init(p->lattice, p->lam_storage, (int)p->n_nodes, rho0, u0,
                                      lam0, du0, drho0,
                                      (hipStream_t)stream),
                  "pylbm_hip_adjoint_init");
}

}  // extern "C"
