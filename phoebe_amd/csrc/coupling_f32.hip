// F32 instantiation of the coupling kernels (coupling_kernel.hpp).
#include "coupling_kernel.hpp"

namespace phx {
template void launch_coupling<float>(const CoupParams&, bool, hipStream_t);
template void launch_separable<float>(const SepParams&, const CoupParams&, bool, int, int, hipStream_t);
template void launch_separable_zy<float>(const SepParams&, int, int, hipStream_t, bool);
template bool launch_separable_x<float>(const SepParams&, const CoupParams&, bool, hipStream_t, hipEvent_t,
                                                  const EpFork&);
int coupling_rows_per_tile_f32(int ldb) { return coupling_rows_per_tile(ldb); }
}  // namespace phx
