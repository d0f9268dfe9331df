// Adjoint, tangent-linear and misfit kernel instantiations for q_u = 1 (n_u = 2),
// both pairings q_v = q_u - 1 and q_v = q_u.
#include "wedg_adjoint.h"

namespace wedg {

hipError_t launch_adj_q1(const AdjArgs& a, int nv, int which, hipStream_t s) {
    if (nv == 2) return launch_adj_nv<2, 2>(a, which, s);
    if (nv == 1) return launch_adj_nv<2, 1>(a, which, s);
    return hipErrorInvalidValue;
}

hipError_t launch_adj2_q1(const AdjArgs& b, const AdjArgs& a, int nv, hipStream_t s) {
    if (nv == 2) return launch_adj2_nv<2, 2>(b, a, s);
    if (nv == 1) return launch_adj2_nv<2, 1>(b, a, s);
    return hipErrorInvalidValue;
}

hipError_t launch_adjb_q1(const BatchArgs& b, int grid, int nv, int which, hipStream_t s) {
    if (nv == 2) return launch_adjb_nv<2, 2>(b, grid, which, s);
    if (nv == 1) return launch_adjb_nv<2, 1>(b, grid, which, s);
    return hipErrorInvalidValue;
}

}  // namespace wedg
