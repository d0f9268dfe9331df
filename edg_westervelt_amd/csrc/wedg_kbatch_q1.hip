// Batched kernel instantiations for q_u = 1 (n_u = 2), both pairings q_v = q_u - 1 and q_v = q_u.
#include "wedg_kernels.h"
#include "wedg_launch.h"
#include "wedg_kern_inst.h"

namespace wedg {

hipError_t launch_batch_q1(const BatchArgs& b, int grid, int nv, int which, hipStream_t s, int mode) {
    if (nv == 2) return launch_batch_nv<2, 2>(b, grid, which, s, mode);
    if (nv == 1) return launch_batch_nv<2, 1>(b, grid, which, s, mode);
    return hipErrorInvalidValue;
}

}  // namespace wedg
