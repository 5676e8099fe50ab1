// Instantiates the adjoint step kernels for D2Q9 with fp16 adjoint states (every
// collision x storage x pattern x force mode, state and record inputs).
#include "lbm_adjoint_kernels.hpp"

namespace plbm {
namespace adj {
template int adj_step_entry<D2Q9, _Float16>(const KArgs&, int, int, int, bool, bool, const void*, const void*, void*, float*, double*, hipStream_t);
}  // namespace adj
}  // namespace plbm
