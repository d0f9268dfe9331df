// Two-lane stage kernel instantiations for q_u = 1 (n_u = 2), both pairings q_v = q_u - 1 and q_v = q_u.
#include "wedg_k2_inst.h"
#include "wedg_launch.h"

namespace wedg {

hipError_t launch2_q1(const KArgs& a, int nv, int which, hipStream_t s, int mode) {
    if (mode != MODE_UNIFORM && mode != MODE_SEPARABLE) return hipErrorInvalidValue;
    if (nv == 2) return launch2_nv<2, 2>(a, which, s, mode);
    if (nv == 1) return launch2_nv<2, 1>(a, which, s, mode);
    return hipErrorInvalidValue;
}

}  // namespace wedg
