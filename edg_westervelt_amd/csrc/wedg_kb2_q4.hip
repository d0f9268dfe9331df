// Batched two-lane stage kernel instantiations for q_u = 4 (n_u = 5), both pairings q_v = q_u - 1 and q_v = q_u.
#include "wedg_k2_inst.h"
#include "wedg_launch.h"

namespace wedg {

hipError_t launch_batch2_q4(const BatchArgs& b, int grid, int nv, int which, hipStream_t s, int mode) {
    if (mode != MODE_UNIFORM && mode != MODE_SEPARABLE) return hipErrorInvalidValue;
    if (nv == 5) return launch_batch2_nv<5, 5>(b, grid, which, s, mode);
    if (nv == 4) return launch_batch2_nv<5, 4>(b, grid, which, s, mode);
    return hipErrorInvalidValue;
}

}  // namespace wedg
