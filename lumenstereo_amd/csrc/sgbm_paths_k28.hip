// sgbm_paths_k28.hip - per-pair aggregation kernels of slices K = 28 and 32.
#include "sgbm_paths_impl.h"

namespace lumen {

template hipError_t launch_paths_k<28>(const PairJob&, const PathsGrid&, int, size_t,
                                      uint16_t*, hipStream_t, StageEvents);
template hipError_t launch_paths_k<32>(const PairJob&, const PathsGrid&, int, size_t,
                                      uint16_t*, hipStream_t, StageEvents);

}  // namespace lumen
