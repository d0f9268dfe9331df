// Adjoint, tangent-linear and misfit kernel instantiations for q_u = 5 (n_u = 6),
// both pairings q_v = q_u - 1 and q_v = q_u.
#include "wedg_adjoint.h"

namespace wedg {

hipError_t launch_adj_q5(const AdjArgs& a, int nv, int which, hipStream_t s) {
    if (nv == 6) return launch_adj_nv<6, 6>(a, which, s);
    if (nv == 5) return launch_adj_nv<6, 5>(a, which, s);
    return hipErrorInvalidValue;
}

hipError_t launch_adj2_q5(const AdjArgs& b, const AdjArgs& a, int nv, hipStream_t s) {
    if (nv == 6) return launch_adj2_nv<6, 6>(b, a, s);
    if (nv == 5) return launch_adj2_nv<6, 5>(b, a, s);
    return hipErrorInvalidValue;
}

hipError_t launch_adjb_q5(const BatchArgs& b, int grid, int nv, int which, hipStream_t s) {
    if (nv == 6) return launch_adjb_nv<6, 6>(b, grid, which, s);
    if (nv == 5) return launch_adjb_nv<6, 5>(b, grid, which, s);
    return hipErrorInvalidValue;
}

}  // namespace wedg
