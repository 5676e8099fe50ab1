// Instantiates the fused step kernels for D2Q9 (BGK / TRT / MRT x 3 storages x 3 patterns
// x 2 force modes).
#include "lbm_step_pair.hpp"

namespace plbm {
template int launch_step<D2Q9>(const KArgs&, int, int, int, int, const void*, void*,
                               hipStream_t);
// two steps per launch on periodic boxes (lbm_step_pair.hpp)
int launch_step_pair_d2q9(const KArgs& a, int coll, int storage, int fmode, const void* src,
                          void* dst, hipStream_t s) {
  return launch_step_pair(a, coll, storage, fmode, src, dst, s);
}
}  // namespace plbm
