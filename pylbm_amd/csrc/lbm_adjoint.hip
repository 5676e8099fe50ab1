// Host-side dispatch of the adjoint kernels (lbm_adjoint_kernels.hpp): the adjoint steps are
// instantiated per lattice and adjoint-state type in adj_<lattice>_<type>.hip; the terminal
// seed and the initial-state gradient are instantiated here.
#include "lbm_adjoint_kernels.hpp"

namespace plbm {
namespace adj {
extern template int adj_step_entry<D2Q9, float>(const KArgs&, int, int, int, bool, bool, const void*,
                                                const void*, void*, float*, double*, hipStream_t);
extern template int adj_step_entry<D2Q9, _Float16>(const KArgs&, int, int, int, bool, bool, const void*,
                                                   const void*, void*, float*, double*, hipStream_t);
extern template int adj_step_entry<D3Q19, float>(const KArgs&, int, int, int, bool, bool, const void*,
                                                 const void*, void*, float*, double*, hipStream_t);
extern template int adj_step_entry<D3Q19, _Float16>(const KArgs&, int, int, int, bool, bool, const void*,
                                                    const void*, void*, float*, double*, hipStream_t);
int adj_pair_d3q19(const KArgs& a, int coll, const void* recA, const void* recB,
                   const void* lin, void* lout, hipStream_t s);
}  // namespace adj

namespace {

using namespace adj;

template <class L, int ST, class LT>
int seed_t(const KArgs& a, bool field, const void* f, const float* G, void* lam, float* dF,
           hipStream_t s) {
  using T = typename Codec<ST>::T;
  const dim3 grid((unsigned)((a.N + kBlock - 1) / kBlock)), block(kBlock);
  if (field)
    k_adjoint_seed<L, ST, true, LT><<<grid, block, 0, s>>>(a, (const T*)f, G, (LT*)lam, dF);
  else
    k_adjoint_seed<L, ST, false, LT><<<grid, block, 0, s>>>(a, (const T*)f, G, (LT*)lam, dF);
  return (int)hipGetLastError();
}

template <class L, class LT>
int seed_l(const KArgs& a, int storage, bool field, const void* f, const float* G, void* lam,
           float* dF, hipStream_t s) {
  switch (storage) {
    case PYLBM_F32: return seed_t<L, PYLBM_F32, LT>(a, field, f, G, lam, dF, s);
    case PYLBM_F16: return seed_t<L, PYLBM_F16, LT>(a, field, f, G, lam, dF, s);
    case PYLBM_BF16: return seed_t<L, PYLBM_BF16, LT>(a, field, f, G, lam, dF, s);
    case PYLBM_F32S: return seed_t<L, PYLBM_F32S, LT>(a, field, f, G, lam, dF, s);
  }
  return PYLBM_EINVAL;
}

template <class L>
int adj_step_lt(const KArgs& a, int coll, int storage, int lam_storage, int pattern, bool field,
                bool rec, const void* f, const void* lin, void* lout, float* dF, double* part,
                hipStream_t s) {
  if (lam_storage == PYLBM_F16)
    return adj_step_entry<L, _Float16>(a, coll, storage, pattern, field, rec, f, lin, lout, dF,
                                       part, s);
  return adj_step_entry<L, float>(a, coll, storage, pattern, field, rec, f, lin, lout, dF,
                                  part, s);
}

template <class L>
int seed_lt(const KArgs& a, int storage, int lam_storage, bool field, const void* f,
            const float* G, void* lam, float* dF, hipStream_t s) {
  if (lam_storage == PYLBM_F16) return seed_l<L, _Float16>(a, storage, field, f, G, lam, dF, s);
  return seed_l<L, float>(a, storage, field, f, G, lam, dF, s);
}

template <class L>
void init_lt(int lam_storage, dim3 grid, int N, const double* rho0, const double* u0,
             const void* lam, double* du, double* drho, hipStream_t s) {
  if (lam_storage == PYLBM_F16)
    k_adjoint_init<L, _Float16><<<grid, dim3(kBlock), 0, s>>>(N, rho0, u0,
                                                              (const _Float16*)lam, du, drho);
  else
    k_adjoint_init<L, float><<<grid, dim3(kBlock), 0, s>>>(N, rho0, u0, (const float*)lam, du,
                                                           drho);
}

}  // namespace

int adjoint_partials(const KArgs& a) { return 2 * a.nblk; }

bool adjoint_pair_applies(const KArgs& a, int lattice, int coll, int lam_storage, bool field) {
  return lattice == PYLBM_D3Q19 && (coll == PYLBM_BGK || coll == PYLBM_TRT) &&
         lam_storage == PYLBM_F32 && !field && a.per == 7 && a.nx <= 128;
}

int launch_adjoint_pair(const KArgs& a, int coll, const void* recA, const void* recB,
                        const void* lin, void* lout, hipStream_t s) {
  return adj::adj_pair_d3q19(a, coll, recA, recB, lin, lout, s);
}

int launch_adjoint_step(int lattice, const KArgs& a, int coll, int storage, int lam_storage,
                        int pattern, bool field, bool rec, const void* f, const void* lin,
                        void* lout, float* dF, double* part, hipStream_t s) {
  if (lattice == PYLBM_D2Q9)
    return adj_step_lt<D2Q9>(a, coll, storage, lam_storage, pattern, field, rec, f, lin, lout,
                             dF, part, s);
  return adj_step_lt<D3Q19>(a, coll, storage, lam_storage, pattern, field, rec, f, lin, lout,
                            dF, part, s);
}

int launch_adjoint_seed(int lattice, const KArgs& a, int storage, int lam_storage, bool field,
                        const void* f, const float* G, void* lam, float* dF, hipStream_t s) {
  if (lattice == PYLBM_D2Q9) return seed_lt<D2Q9>(a, storage, lam_storage, field, f, G, lam, dF, s);
  return seed_lt<D3Q19>(a, storage, lam_storage, field, f, G, lam, dF, s);
}

int launch_adjoint_init(int lattice, int lam_storage, int N, const double* rho0,
                        const double* u0, const void* lam, double* du, double* drho,
                        hipStream_t s) {
  const dim3 grid((unsigned)((N + kBlock - 1) / kBlock));
  if (lattice == PYLBM_D2Q9)
    init_lt<D2Q9>(lam_storage, grid, N, rho0, u0, lam, du, drho, s);
  else
    init_lt<D3Q19>(lam_storage, grid, N, rho0, u0, lam, du, drho, s);
  return (int)hipGetLastError();
}

}  // namespace plbm
