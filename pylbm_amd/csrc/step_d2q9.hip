// Instantiates the fused step kernels for D2Q9 (BGK / TRT / MRT x 3 storages x 3 patterns
// x 2 force modes).
#include "lbm_step_pair.hpp"

namespace plbm {
template int launch_step<D2Q9>(const KArgs&, int, int, int, int, const void*, void*,
                               hipStream_t);
// two steps per launch on box problems (lbm_step_pair.hpp)
int launch_step_pair_d2q9(const KArgs& a, int coll, int storage, int fmode, const void* src,
                          void* dst, hipStream_t s) {
  return launch_step_pair(a, coll, storage, fmode, src, dst, s);
}
// the access pattern of the kernel above without its arithmetic (pylbm_hip_sweep_probe)
int launch_sweep_probe_d2q9(const KArgs& a, int elem_bytes, const void* src, void* dst,
                            hipStream_t s) {
  return elem_bytes == 4 ? launch_sweep_probe_t<PYLBM_F32>(a, src, dst, s)
                         : launch_sweep_probe_t<PYLBM_F16>(a, src, dst, s);
}
}  // namespace plbm
