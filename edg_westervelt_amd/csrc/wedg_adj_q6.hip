// Adjoint, tangent-linear and misfit kernel instantiations for q_u = 6 (n_u = 7),
// both pairings q_v = q_u - 1 and q_v = q_u.
#include "wedg_adjoint.h"

namespace wedg {

hipError_t launch_adj_q6(const AdjArgs& a, int nv, int which, hipStream_t s) {
    if (nv == 7) return launch_adj_nv<7, 7>(a, which, s);
    if (nv == 6) return launch_adj_nv<7, 6>(a, which, s);
    return hipErrorInvalidValue;
}

hipError_t launch_adj2_q6(const AdjArgs& b, const AdjArgs& a, int nv, hipStream_t s) {
    if (nv == 7) return launch_adj2_nv<7, 7>(b, a, s);
    if (nv == 6) return launch_adj2_nv<7, 6>(b, a, s);
    return hipErrorInvalidValue;
}

hipError_t launch_adjb_q6(const BatchArgs& b, int grid, int nv, int which, hipStream_t s) {
    if (nv == 7) return launch_adjb_nv<7, 7>(b, grid, which, s);
    if (nv == 6) return launch_adjb_nv<7, 6>(b, grid, which, s);
    return hipErrorInvalidValue;
}

}  // namespace wedg
