// sgbm_paths_k20.hip - per-pair aggregation kernels of slices K = 20 and 24.
#include "sgbm_paths_impl.h"

namespace lumen {

template hipError_t launch_paths_k<20>(const PairJob&, const PathsGrid&, int, size_t,
                                      uint16_t*, hipStream_t, StageEvents);
template hipError_t launch_paths_k<24>(const PairJob&, const PathsGrid&, int, size_t,
                                      uint16_t*, hipStream_t, StageEvents);

}  // namespace lumen
