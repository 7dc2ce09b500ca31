// force_f32.hip -- FP32-arithmetic instantiations of the force step: FP32
// positions, and FP64 positions differenced before the FP32 force math.
#include "force_typed.hpp"

namespace gvh {
template int typed_step<float, float>(ForceCommon &, int);
template int typed_step<float, double>(ForceCommon &, int);
template int typed_upload<float>(ForceCommon &);
template int typed_batch_step<float, float>(ForceCommon &, BatchState &);
template int typed_batch_step<float, double>(ForceCommon &, BatchState &);
} // namespace gvh
