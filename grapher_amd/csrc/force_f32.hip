// force_f32.hip -- FP32-arithmetic instantiations of the force step: FP32
// positions, and FP64 positions differenced before the FP32 force math.
#include "force_typed.hpp"

namespace gvh {
template int typed_step<float, float>(ForceCommon &, int);
template int typed_step<float, double>(ForceCommon &, int);
template int typed_upload<float>(ForceCommon &);
} // namespace gvh
