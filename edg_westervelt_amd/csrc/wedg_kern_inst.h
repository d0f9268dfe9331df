// The kernel launches of one (n_u, n_v) pair, included by the per-degree
// instantiation files wedg_kern_q*.hip.
#pragma once
#include <hip/hip_runtime.h>
#include "wedg_kernels.h"

namespace wedg {

// The launches of one (n_u, n_v) pair; used by the instantiation files only.
template <int NU, int NV>
static hipError_t launch_nv(const KArgs& a, int which, hipStream_t s, int mode) {
    const int grid = (a.E + kWave - 1) / kWave;
    if (which == 0) {
        // the problem's own record layout: class records (uniform, separable)
        // or per-element element-fastest ones (general, conjugate gradients)
        if (mode == MODE_UNIFORM) k_prep<NU, NV, MODE_UNIFORM><<<grid, kWave, 0, s>>>(a);
        else if (mode == MODE_SEPARABLE) k_prep<NU, NV, MODE_SEPARABLE><<<grid, kWave, 0, s>>>(a);
        else k_prep<NU, NV, MODE_GENERAL><<<grid, kWave, 0, s>>>(a);
    } else if (which == 1) {
        if (mode == MODE_UNIFORM) k_stage<NU, NV, MODE_UNIFORM><<<grid, kWave, 0, s>>>(a);
        else if (mode == MODE_SEPARABLE) k_stage<NU, NV, MODE_SEPARABLE><<<grid, kWave, 0, s>>>(a);
        else if (mode == MODE_GENERAL_CG) k_stage<NU, NV, MODE_GENERAL_CG><<<grid, kWave, 0, s>>>(a);
        else k_stage<NU, NV, MODE_GENERAL><<<grid, kWave, 0, s>>>(a);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// The batched launches of one (n_u, n_v) pair: grid blocks over the instances
// of b; used by the batched instantiation files only.
template <int NU, int NV>
static hipError_t launch_batch_nv(const BatchArgs& b, int grid, int which, hipStream_t s, int mode) {
    if (which == 0) {
        if (mode == MODE_UNIFORM) k_prep<NU, NV, MODE_UNIFORM, true><<<grid, kWave, 0, s>>>(b);
        else if (mode == MODE_SEPARABLE) k_prep<NU, NV, MODE_SEPARABLE, true><<<grid, kWave, 0, s>>>(b);
        else k_prep<NU, NV, MODE_GENERAL, true><<<grid, kWave, 0, s>>>(b);
    } else if (which == 1) {
        if (mode == MODE_UNIFORM) k_stage<NU, NV, MODE_UNIFORM, true><<<grid, kWave, 0, s>>>(b);
        else if (mode == MODE_SEPARABLE) k_stage<NU, NV, MODE_SEPARABLE, true><<<grid, kWave, 0, s>>>(b);
        else if (mode == MODE_GENERAL_CG) k_stage<NU, NV, MODE_GENERAL_CG, true><<<grid, kWave, 0, s>>>(b);
        else k_stage<NU, NV, MODE_GENERAL, true><<<grid, kWave, 0, s>>>(b);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int NU, int NV>
static hipError_t launch_diag_nv(const DiagArgs* d, const RecvArgs* r, const NormArgs* n, hipStream_t s) {
    if (d) k_diag<NU, NV><<<(d->E + kWave - 1) / kWave, kWave, 0, s>>>(*d);
    if (r) k_recv<NV><<<(r->nrecv + kWave - 1) / kWave, kWave, 0, s>>>(*r);
    if (n) k_norms<NU, NV><<<(n->E + kWave - 1) / kWave, kWave, 0, s>>>(*n);
    return hipGetLastError();
}

}  // namespace wedg
