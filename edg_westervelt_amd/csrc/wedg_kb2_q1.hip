// Batched two-lane stage kernel instantiations for q_u = 1 (n_u = 2), both pairings q_v = 0 and q_v = 1.
#include "wedg_k2_inst.h"
#include "wedg_launch.h"

namespace wedg {

hipError_t launch_batch2_q1(const BatchArgs& b, int grid, int nv, int which, hipStream_t s, int mode) {
    if (mode != MODE_UNIFORM && mode != MODE_SEPARABLE) return hipErrorInvalidValue;
    if (nv == 2) return launch_batch2_nv<2, 2>(b, grid, which, s, mode);
    if (nv == 1) return launch_batch2_nv<2, 1>(b, grid, which, s, mode);
    return hipErrorInvalidValue;
}

}  // namespace wedg
