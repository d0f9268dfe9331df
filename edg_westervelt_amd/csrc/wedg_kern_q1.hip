// Kernel instantiations for q_u = 1 (n_u = 2), both pairings q_v = q_u - 1 and q_v = q_u.
#include "wedg_kernels.h"
#include "wedg_launch.h"
#include "wedg_kern_inst.h"

namespace wedg {

hipError_t launch_q1(const KArgs& a, int nv, int which, hipStream_t s, int mode) {
    if (nv == 2) return launch_nv<2, 2>(a, which, s, mode);
    if (nv == 1) return launch_nv<2, 1>(a, which, s, mode);
    return hipErrorInvalidValue;
}

hipError_t launch_diag_q1(const DiagArgs* d, const RecvArgs* r, const NormArgs* n, int nv, hipStream_t s) {
    if (nv == 2) return launch_diag_nv<2, 2>(d, r, n, s);
    if (nv == 1) return launch_diag_nv<2, 1>(d, r, n, s);
    return hipErrorInvalidValue;
}

}  // namespace wedg
