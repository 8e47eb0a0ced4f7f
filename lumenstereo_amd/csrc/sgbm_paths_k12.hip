// sgbm_paths_k12.hip - per-pair aggregation kernels of slices K = 12 and 16.
#include "sgbm_paths_impl.h"

namespace lumen {

template hipError_t launch_paths_k<12>(const PairJob&, const PathsGrid&, int, size_t,
                                      uint16_t*, hipStream_t, StageEvents);
template hipError_t launch_paths_k<16>(const PairJob&, const PathsGrid&, int, size_t,
                                      uint16_t*, hipStream_t, StageEvents);

}  // namespace lumen
