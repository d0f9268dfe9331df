// wedg_rhs_noalloc (include/wedg.h): the right-hand side of wedg_rhs_device
// without a workspace, for callers that cannot allocate device memory. It has
// no counterpart in the reference.
//
// Besides the state and the problem's tables, the stage kernels need scratch:
// the boundary face slots [side][i][A, D, F][N], the failure record and the
// per-wave Chebyshev counts and count statistics. A workspace allocates them;
// here they live in one static device array, `scratch`, of
// WEDG_NOALLOC_SCRATCH_BYTES, which every GPU the library runs on holds until
// the process ends. Its size bounds the mesh at N <= WEDG_NOALLOC_MAX_N (q_u <= 6).
//
// In host code `scratch` is a host-side stand-in, not a device address: each
// GPU holds its own copy at its own address, looked up once per device with
// hipGetSymbolAddress and kept in g_scratch; only that address reaches the
// kernels. Nothing here allocates: the clearing of the record and of the count
// statistics is a small kernel (no hipMemsetAsync), and the failure record is
// mapped on the device into the caller's wedg_error_info.
#include "wedg_internal.h"

#include <mutex>

namespace {

constexpr int kMaxN = WEDG_NOALLOC_MAX_N;
constexpr int kMaxNU = 7;   // q_u <= 6
constexpr size_t kMaxWaves = ((size_t)kMaxN * kMaxN + kWave - 1) / kWave;
constexpr size_t kOffErr = 0;                                                       // DevErr
constexpr size_t kOffBsl = 256;                                                     // face slots
constexpr size_t kOffCnt = kOffBsl + (size_t)4 * kMaxNU * 3 * kMaxN * sizeof(double);   // counts [2][waves]
constexpr size_t kOffStat = kOffCnt + 2 * kMaxWaves * sizeof(int);                  // statistics [2][waves]
constexpr size_t kScratchBytes = kOffStat + 2 * kMaxWaves * sizeof(unsigned);
static_assert(sizeof(DevErr) <= kOffBsl, "the failure record fits its slot");
static_assert(kOffCnt % 256 == 0 && kOffStat % 256 == 0, "aligned regions");
static_assert(kScratchBytes == WEDG_NOALLOC_SCRATCH_BYTES, "the size include/wedg.h states");

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up on the
// first call on that device.
wedg_status scratch_of_current_device(unsigned char** out) {
    int dev = -1;
    HIPOK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return WEDG_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        HIPOK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return WEDG_ERR_HIP;
        g_scratch[dev] = static_cast<unsigned char*>(ptr);
    }
    *out = g_scratch[dev];
    return WEDG_OK;
}

// The failure record as a fresh call starts it, and zero count statistics.
__global__ void k_noalloc_clear(DevErr* err, unsigned* wstat, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        DevErr z = {};
        z.fkey = ~0ull;
        z.nonfinite = INT_MAX;
        z.nf_serial = INT_MAX;
        *err = z;
    }
    for (int k = i; k < n; k += gridDim.x * blockDim.x) wstat[k] = 0u;
}

// The record mapped to the caller's wedg_error_info, with the largest count.
__global__ __launch_bounds__(256) void k_noalloc_report(const DevErr* err, const unsigned* wstat, int nw, int nq,
                                                        int N, wedg_error_info* info) {
    __shared__ unsigned red[256];
    unsigned m = 0u;
    for (int w = threadIdx.x; w < nw; w += blockDim.x) m = max(m, wstat[nw + w]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const DevErr h = *err;
        wedg_error_info L = {};
        L.cheb_steps = (int)red[0];
        L.cg_unconverged = h.cg_unconverged;
        map_failure(h, nq, N, 0, L);
        *info = L;
    }
}

}  // namespace

extern "C" wedg_status wedg_rhs_noalloc(const wedg_problem* p, int cheb_policy, const double* u, const double* v,
                                        double t, double* ut, double* vt, wedg_error_info* info, void* stream) {
    if (!p || !u || !v || !ut || !vt) return WEDG_ERR_ARGUMENT;
    if (cheb_policy != WEDG_CHEB_GLOBAL && cheb_policy != WEDG_CHEB_ELEMENT) return WEDG_ERR_ARGUMENT;
    if (p->N > kMaxN || p->nu > kMaxNU) return WEDG_ERR_SCRATCH;
    unsigned char* base = nullptr;
    wedg_status s;
    if ((s = scratch_of_current_device(&base))) return s;
    hipStream_t st = static_cast<hipStream_t>(stream);
    DevErr* err = reinterpret_cast<DevErr*>(base + kOffErr);
    double* bsl = reinterpret_cast<double*>(base + kOffBsl);
    int* cnt = reinterpret_cast<int*>(base + kOffCnt);
    unsigned* wstat = reinterpret_cast<unsigned*>(base + kOffStat);
    const int nw = (int)(((size_t)p->E + kWave - 1) / kWave);

    const int grid = std::min(64, (2 * nw + 255) / 256);
    k_noalloc_clear<<<grid, 256, 0, st>>>(err, wstat, 2 * nw);
    HIPOK(hipGetLastError());
    // stage 0 (a plain right-hand side) with the workspace's buffers replaced by
    // the scratch regions; k_prep forms the counts the stage reads (slot 1)
    KArgs a = problem_args(p, cheb_policy == WEDG_CHEB_GLOBAL, bsl, err, cnt, wstat);
    a.stage = 0;
    a.serial = 1;
    a.slot = 1;
    a.step = 0;
    a.cstage = 0.0;
    stage_time(p, t, a);
    a.yu = u; a.yv = v; a.y0u = u; a.y0v = v;
    a.outu = ut; a.outv = vt;
    a.has_next = 0;
    wedg_problem* pm = const_cast<wedg_problem*>(p);
    HIPOK(launch(a, pm, 0, st, false));   // one lane per element: the scratch holds 64-element waves
    HIPOK(launch(a, pm, 1, st, false));
    if (info) {
        k_noalloc_report<<<1, 256, 0, st>>>(err, wstat, nw, p->nu, p->N, info);
        HIPOK(hipGetLastError());
    }
    return WEDG_OK;
}
