// Kernel instantiations for q_u = 6 (n_u = 7), both pairings q_v = q_u - 1 and q_v = q_u.
#include "wedg_kernels.h"
#include "wedg_launch.h"

namespace wedg {

template <int NU, int NV>
static hipError_t launch_nv(const KArgs& a, int which, hipStream_t s, int uni) {
    const int grid = (a.E + kBlock - 1) / kBlock;
    switch (which) {
        case 0: k_traces<NU, NV><<<grid, kBlock, 0, s>>>(a, a.tab, a.beta); break;
        case 1:
            if (uni) k_volume<NU, NV, true><<<grid, kBlock, 0, s>>>(a, a.tab, a.Cw, a.Aw, a.Rbuf, a.Bbuf);
            else k_volume<NU, NV, false><<<grid, kBlock, 0, s>>>(a, a.tab, a.Cw, a.Aw, a.Rbuf, a.Bbuf);
            break;
        case 2:
            if (uni) k_faces<NU, NV, true><<<grid, kBlock, 0, s>>>(a, a.tab, a.Ainv, a.Rbuf, a.Bbuf);
            else k_faces<NU, NV, false><<<grid, kBlock, 0, s>>>(a, a.tab, a.Ainv, a.Rbuf, a.Bbuf);
            break;
        case 3:
            k_mass<NU, NV><<<(a.E + kMassBlock - 1) / kMassBlock, kMassBlock, 0, s>>>(a, a.tab, a.beta, a.Rbuf);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_q6(const KArgs& a, int nv, int which, hipStream_t s, int uni) {
    if (nv == 7) return launch_nv<7, 7>(a, which, s, uni);
    if (nv == 6) return launch_nv<7, 6>(a, which, s, uni);
    return hipErrorInvalidValue;
}

}  // namespace wedg
