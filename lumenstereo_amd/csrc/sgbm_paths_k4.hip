// sgbm_paths_k4.hip - per-pair aggregation kernels of slices K = 4, 6 and 8.
#include "sgbm_paths_impl.h"

namespace lumen {

template hipError_t launch_paths_k<4>(const PairJob&, const PathsGrid&, int, size_t,
                                      uint16_t*, hipStream_t, StageEvents);
template hipError_t launch_paths_k<6>(const PairJob&, const PathsGrid&, int, size_t,
                                      uint16_t*, hipStream_t, StageEvents);
template hipError_t launch_paths_k<8>(const PairJob&, const PathsGrid&, int, size_t,
                                      uint16_t*, hipStream_t, StageEvents);

}  // namespace lumen
