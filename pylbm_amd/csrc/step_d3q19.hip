// Instantiates the fused step kernels for D3Q19 (BGK / TRT x 3 storages x 3 patterns x
// 2 force modes).
#include "lbm_step.hpp"

namespace plbm {
template int launch_step<D3Q19>(const KArgs&, int, int, int, int, const void*, void*,
                                hipStream_t);
}  // namespace plbm
