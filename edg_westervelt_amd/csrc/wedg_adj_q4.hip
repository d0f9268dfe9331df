// Adjoint, tangent-linear and misfit kernel instantiations for q_u = 4 (n_u = 5),
// both pairings q_v = q_u - 1 and q_v = q_u.
#include "wedg_adjoint.h"

namespace wedg {

hipError_t launch_adj_q4(const AdjArgs& a, int nv, int which, hipStream_t s) {
    if (nv == 5) return launch_adj_nv<5, 5>(a, which, s);
    if (nv == 4) return launch_adj_nv<5, 4>(a, which, s);
    return hipErrorInvalidValue;
}

hipError_t launch_adj2_q4(const AdjArgs& b, const AdjArgs& a, int nv, hipStream_t s) {
    if (nv == 5) return launch_adj2_nv<5, 5>(b, a, s);
    if (nv == 4) return launch_adj2_nv<5, 4>(b, a, s);
    return hipErrorInvalidValue;
}

hipError_t launch_adjb_q4(const BatchArgs& b, int grid, int nv, int which, hipStream_t s) {
    if (nv == 5) return launch_adjb_nv<5, 5>(b, grid, which, s);
    if (nv == 4) return launch_adjb_nv<5, 4>(b, grid, which, s);
    return hipErrorInvalidValue;
}

}  // namespace wedg
