// Instantiates the general-geometry step kernel and the link-mask kernels for D3Q19.
#include "lbm_geometry.hpp"

namespace plbm {
template int launch_step_geo<D3Q19>(const KArgs&, int, int, int, const void*, void*, hipStream_t);
template int launch_links<D3Q19>(const uint8_t*, void*, int, int, int, hipStream_t);
template int launch_link_groups<D3Q19>(const void*, uint32_t*, int, int, int, hipStream_t);
}  // namespace plbm
