// Tile kernels (single mesh and batched) of the FP64 form on Q3 meshes: every
// variant of every mode, in a translation unit of their own so that the kernels
// compile in parallel.
#include "tile_kernel.hpp"

namespace lbfem_amd
{
  template hipError_t tile_launch_degree<double, 3>(int, int, const Geo &, const TileParams &, const double *,
                                                     const double *, double *, double *, hipStream_t);
  template hipError_t tile_batch_launch_degree<double, 3>(int, int, const BatchEntry<double> *, const BatchCounts *, int, int,
                                                           hipStream_t);
} // namespace lbfem_amd
