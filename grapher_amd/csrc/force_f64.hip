// force_f64.hip -- FP64 instantiation of the force step. Built with
// -ffp-contract=off so every product and sum rounds as in the reference.
#include "force_typed.hpp"

namespace gvh {
template int typed_step<double, double>(ForceCommon &, int);
template int typed_upload<double>(ForceCommon &);
template int typed_batch_step<double, double>(ForceCommon &, BatchState &);
} // namespace gvh
