// Batched kernel instantiations for q_u = 2 (n_u = 3), both pairings q_v = q_u - 1 and q_v = q_u.
#include "wedg_kernels.h"
#include "wedg_launch.h"
#include "wedg_kern_inst.h"

namespace wedg {

hipError_t launch_batch_q2(const BatchArgs& b, int grid, int nv, int which, hipStream_t s, int mode) {
    if (nv == 3) return launch_batch_nv<3, 3>(b, grid, which, s, mode);
    if (nv == 2) return launch_batch_nv<3, 2>(b, grid, which, s, mode);
    return hipErrorInvalidValue;
}

}  // namespace wedg
