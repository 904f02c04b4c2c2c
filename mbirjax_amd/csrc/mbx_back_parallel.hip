// Back projector kernels of the parallel geometry (one translation unit per
// geometry, so that their instantiations compile in parallel).
#include "mbx_back_kernels.hpp"

namespace mbx {
template int dispatch_back_1<MBX_GEOM_PARALLEL>(const DevGeom &, const BackPlan &, const float *, const void *, int,
                                 const float *, int, const int32_t *, int, int, float, int, float *, float *,
                                 hipStream_t);
}  // namespace mbx
