// Kernel instantiations for q_u = 4 (n_u = 5), both pairings q_v = q_u - 1 and q_v = q_u.
#include "wedg_kernels.h"
#include "wedg_launch.h"
#include "wedg_kern_inst.h"

namespace wedg {

hipError_t launch_q4(const KArgs& a, int nv, int which, hipStream_t s, int mode) {
    if (nv == 5) return launch_nv<5, 5>(a, which, s, mode);
    if (nv == 4) return launch_nv<5, 4>(a, which, s, mode);
    return hipErrorInvalidValue;
}

hipError_t launch_diag_q4(const DiagArgs* d, const RecvArgs* r, const NormArgs* n, int nv, hipStream_t s) {
    if (nv == 5) return launch_diag_nv<5, 5>(d, r, n, s);
    if (nv == 4) return launch_diag_nv<5, 4>(d, r, n, s);
    return hipErrorInvalidValue;
}

}  // namespace wedg
