// Instantiates the general-geometry step kernel and the link-mask kernels for D2Q9.
#include "lbm_geometry.hpp"

namespace plbm {
template int launch_step_geo<D2Q9>(const KArgs&, int, int, int, const void*, void*, hipStream_t);
template int launch_links<D2Q9>(const uint8_t*, void*, int, int, int, hipStream_t);
template int launch_link_groups<D2Q9>(const void*, uint32_t*, int, int, int, hipStream_t);
}  // namespace plbm
