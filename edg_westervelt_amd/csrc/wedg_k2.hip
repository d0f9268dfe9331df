// Instantiations of the two-lane stage kernels (wedg_stage2.h): q_u = q_v odd
// (n_u = n_v = 2, 4, 6), the spectral u-solve modes.
#include "wedg_stage2.h"
#include "wedg_launch.h"

namespace wedg {

template <int NU>
static hipError_t launch2_nu(const KArgs& a, int which, hipStream_t s, int mode) {
    const int grid = (a.E + kPairWave - 1) / kPairWave;
    if (which == 0) {
        if (mode == MODE_UNIFORM) k_prep2<NU, NU, MODE_UNIFORM><<<grid, kWave, 0, s>>>(a);
        else k_prep2<NU, NU, MODE_SEPARABLE><<<grid, kWave, 0, s>>>(a);
    } else if (which == 1) {
        if (mode == MODE_UNIFORM) k_stage2<NU, NU, MODE_UNIFORM><<<grid, kWave, 0, s>>>(a);
        else if (mode == MODE_SEPARABLE) k_stage2<NU, NU, MODE_SEPARABLE><<<grid, kWave, 0, s>>>(a);
        else return hipErrorInvalidValue;
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch2(const KArgs& a, int nu, int nv, int which, hipStream_t s, int mode) {
    if (nv != nu || (mode != MODE_UNIFORM && mode != MODE_SEPARABLE)) return hipErrorInvalidValue;
    switch (nu) {
        case 2: return launch2_nu<2>(a, which, s, mode);
        case 4: return launch2_nu<4>(a, which, s, mode);
        case 6: return launch2_nu<6>(a, which, s, mode);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace wedg
