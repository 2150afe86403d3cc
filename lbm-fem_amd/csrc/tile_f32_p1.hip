// Tile kernels (single mesh and batched) of the mixed form (FP32 deviations) on
// Q1 meshes: every variant of every mode, in a translation unit of their own so
// that the kernels compile in parallel.
#include "tile_kernel.hpp"

namespace lbfem_amd
{
  template hipError_t tile_launch_degree<float, 1>(int, int, const Geo &, const TileParams &, const float *,
                                                     const float *, float *, float *, hipStream_t);
  template hipError_t tile_batch_launch_degree<float, 1>(int, int, const BatchEntry<float> *, const BatchCounts *, int, int,
                                                           hipStream_t);
} // namespace lbfem_amd
