// Launch entry points of the per-degree kernel instantiations.
#pragma once
#include <hip/hip_runtime.h>
#include "wedg_kernels.h"

namespace wedg {

// which: 0 = k_prep, 1 = k_stage; mode: MODE_GENERAL, MODE_SEPARABLE, MODE_UNIFORM or MODE_GENERAL_CG
hipError_t launch_q1(const KArgs& a, int nv, int which, hipStream_t s, int mode);
hipError_t launch_q2(const KArgs& a, int nv, int which, hipStream_t s, int mode);
hipError_t launch_q3(const KArgs& a, int nv, int which, hipStream_t s, int mode);
hipError_t launch_q4(const KArgs& a, int nv, int which, hipStream_t s, int mode);
hipError_t launch_q5(const KArgs& a, int nv, int which, hipStream_t s, int mode);
hipError_t launch_q6(const KArgs& a, int nv, int which, hipStream_t s, int mode);

// batched launches: grid blocks over the instances of b (which and mode as above)
hipError_t launch_batch_q1(const BatchArgs& b, int grid, int nv, int which, hipStream_t s, int mode);
hipError_t launch_batch_q2(const BatchArgs& b, int grid, int nv, int which, hipStream_t s, int mode);
hipError_t launch_batch_q3(const BatchArgs& b, int grid, int nv, int which, hipStream_t s, int mode);
hipError_t launch_batch_q4(const BatchArgs& b, int grid, int nv, int which, hipStream_t s, int mode);
hipError_t launch_batch_q5(const BatchArgs& b, int grid, int nv, int which, hipStream_t s, int mode);
hipError_t launch_batch_q6(const BatchArgs& b, int grid, int nv, int which, hipStream_t s, int mode);

// the two-lane stage kernels (wedg_stage2.h; which as above, MODE_UNIFORM or MODE_SEPARABLE)
hipError_t launch2_q1(const KArgs& a, int nv, int which, hipStream_t s, int mode);
hipError_t launch2_q2(const KArgs& a, int nv, int which, hipStream_t s, int mode);
hipError_t launch2_q3(const KArgs& a, int nv, int which, hipStream_t s, int mode);
hipError_t launch2_q4(const KArgs& a, int nv, int which, hipStream_t s, int mode);
hipError_t launch2_q5(const KArgs& a, int nv, int which, hipStream_t s, int mode);
hipError_t launch2_q6(const KArgs& a, int nv, int which, hipStream_t s, int mode);
hipError_t launch_batch2_q1(const BatchArgs& b, int grid, int nv, int which, hipStream_t s, int mode);
hipError_t launch_batch2_q2(const BatchArgs& b, int grid, int nv, int which, hipStream_t s, int mode);
hipError_t launch_batch2_q3(const BatchArgs& b, int grid, int nv, int which, hipStream_t s, int mode);
hipError_t launch_batch2_q4(const BatchArgs& b, int grid, int nv, int which, hipStream_t s, int mode);
hipError_t launch_batch2_q5(const BatchArgs& b, int grid, int nv, int which, hipStream_t s, int mode);
hipError_t launch_batch2_q6(const BatchArgs& b, int grid, int nv, int which, hipStream_t s, int mode);

// the diagnostics and norm kernels (any argument may be null)
hipError_t launch_diag_q1(const DiagArgs* d, const RecvArgs* r, const NormArgs* n, int nv, hipStream_t s);
hipError_t launch_diag_q2(const DiagArgs* d, const RecvArgs* r, const NormArgs* n, int nv, hipStream_t s);
hipError_t launch_diag_q3(const DiagArgs* d, const RecvArgs* r, const NormArgs* n, int nv, hipStream_t s);
hipError_t launch_diag_q4(const DiagArgs* d, const RecvArgs* r, const NormArgs* n, int nv, hipStream_t s);
hipError_t launch_diag_q5(const DiagArgs* d, const RecvArgs* r, const NormArgs* n, int nv, hipStream_t s);
hipError_t launch_diag_q6(const DiagArgs* d, const RecvArgs* r, const NormArgs* n, int nv, hipStream_t s);

}  // namespace wedg
