// Instantiates the adjoint step kernels for D3Q19 with fp32 adjoint states (every
// collision x storage x pattern x force mode, state and record inputs), and the two-step record kernel.
#include "lbm_adjoint_pair.hpp"

namespace plbm {
namespace adj {
template int adj_step_entry<D3Q19, float>(const KArgs&, int, int, int, bool, bool, const void*, const void*, void*, float*, double*, hipStream_t);

// Two record steps per launch (lbm_adjoint_pair.hpp); PYLBM_ENOTSUP when they do not apply.
int adj_pair_d3q19(const KArgs& a, int coll, const void* recA, const void* recB,
                   const void* lin, void* lout, hipStream_t s) {
  if (coll == PYLBM_BGK) return launch_pair_c<PYLBM_BGK>(a, recA, recB, lin, lout, s);
  if (coll == PYLBM_TRT) return launch_pair_c<PYLBM_TRT>(a, recA, recB, lin, lout, s);
  return PYLBM_ENOTSUP;
}
}  // namespace adj
}  // namespace plbm
