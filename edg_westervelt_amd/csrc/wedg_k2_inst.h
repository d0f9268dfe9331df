// The launches of the two-lane stage kernels (wedg_stage2.h) of one (n_u, n_v)
// pair; used by the instantiation files wedg_k2_q*.hip only.
#pragma once
#include <hip/hip_runtime.h>
#include "wedg_stage2.h"

namespace wedg {

template <int NU, int NV>
static hipError_t launch2_nv(const KArgs& a, int which, hipStream_t s, int mode) {
    const int grid = (a.E + kPairWave - 1) / kPairWave;
    if (which == 0) {
        if (mode == MODE_UNIFORM) k_prep2<NU, NV, MODE_UNIFORM><<<grid, kWave, 0, s>>>(a);
        else k_prep2<NU, NV, MODE_SEPARABLE><<<grid, kWave, 0, s>>>(a);
    } else if (which == 1) {
        if (mode == MODE_UNIFORM) k_stage2<NU, NV, MODE_UNIFORM><<<grid, kWave, 0, s>>>(a);
        else if (mode == MODE_SEPARABLE) k_stage2<NU, NV, MODE_SEPARABLE><<<grid, kWave, 0, s>>>(a);
        else return hipErrorInvalidValue;
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// The batched launches (grid blocks over the instances of b, each a range of
// waves of 32 elements); used by the instantiation files wedg_kb2_q*.hip only.
template <int NU, int NV>
static hipError_t launch_batch2_nv(const BatchArgs& b, int grid, int which, hipStream_t s, int mode) {
    if (which == 0) {
        if (mode == MODE_UNIFORM) k_prep2<NU, NV, MODE_UNIFORM, true><<<grid, kWave, 0, s>>>(b);
        else k_prep2<NU, NV, MODE_SEPARABLE, true><<<grid, kWave, 0, s>>>(b);
    } else if (which == 1) {
        if (mode == MODE_UNIFORM) k_stage2<NU, NV, MODE_UNIFORM, true><<<grid, kWave, 0, s>>>(b);
        else if (mode == MODE_SEPARABLE) k_stage2<NU, NV, MODE_SEPARABLE, true><<<grid, kWave, 0, s>>>(b);
        else return hipErrorInvalidValue;
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace wedg
