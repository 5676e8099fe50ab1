// Instantiates the fused step kernels for D2Q9 (BGK / TRT / MRT x 3 storages x 3 patterns
// x 2 force modes).
#include "lbm_step.hpp"

namespace plbm {
template int launch_step<D2Q9>(const KArgs&, int, int, int, int, const void*, void*,
                               hipStream_t);
}  // namespace plbm
