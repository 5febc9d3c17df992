// Batched window kernel (cells per workgroup, node state staged in LDS).
#include <hip/hip_runtime.h>

#include "layout.h"

namespace mind {
struct State;
int batched_setup(State&, int32_t) { return MIND_EUNSUPPORTED; }
int batched_launch(State&, int32_t) { return MIND_EUNSUPPORTED; }
}  // namespace mind
