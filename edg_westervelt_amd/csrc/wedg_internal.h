// Internal declarations shared by the host translation units of libwedg.so
// (wedg_host.hip: the forward C ABI; wedg_adjoint.hip: the discrete adjoint):
// the problem and workspace structures and the stage helpers. Not part of the
// public ABI (include/wedg.h).
#pragma once
#include "../../include/wedg.h"
#include "wedg_kernels.h"
#include "wedg_launch.h"
#include "wedg_split.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#define HIPOK(x)                                   \
    do {                                           \
        hipError_t e_ = (x);                       \
        if (e_ != hipSuccess) return WEDG_ERR_HIP; \
    } while (0)

namespace wedg {

template <class T>
T* dalloc(size_t n, hipError_t* err) {
    void* p = nullptr;
    *err = hipMalloc(&p, n * sizeof(T) + (n == 0 ? 8 : 0));
    return static_cast<T*>(p);
}

uint64_t fnv1a(const double* v, size_t n);

// Classes of identical per-element records (bitwise). rec(e, buf) fills the
// record of element e; returns the class per element and the unique records.
template <class F>
void dedup(long E, size_t len, F rec, std::vector<int>& cls, std::vector<double>& uniq) {
    std::unordered_map<uint64_t, std::vector<int>> map;
    std::vector<double> buf(len);
    cls.assign(E, 0);
    uniq.clear();
    int ncls = 0;
    for (long e = 0; e < E; ++e) {
        rec(e, buf.data());
        uint64_t h = fnv1a(buf.data(), len);
        auto& cands = map[h];
        int found = -1;
        for (int c : cands)
            if (std::memcmp(&uniq[(size_t)c * len], buf.data(), len * sizeof(double)) == 0) { found = c; break; }
        if (found < 0) {
            found = ncls++;
            uniq.insert(uniq.end(), buf.begin(), buf.end());
            cands.push_back(found);
        }
        cls[e] = found;
    }
}

// The failure record of a run mapped to the error it reports: the first
// failure in serial order, a boundary weight before the mass matrix of the same
// stage (src/rhs.jl:530-531). A non-finite state is an error only when asked
// (the blow-up check of src/solver.jl:940, after the step): then it comes
// before every failure of a later stage, which the device records as it runs
// on from the non-finite state. Sets status, key, side, element and step of L;
// returns the status.
__host__ __device__ inline int map_failure(const DevErr& h, int nq, int N, int nonfinite_is_error,
                                           wedg_error_info& L) {
    const bool blown = nonfinite_is_error && h.nonfinite != INT_MAX;
    if (h.fkey != ~0ull && !(blown && (long long)(h.fkey >> 33) > (long long)h.nf_serial)) {
        const int kind = (int)((h.fkey >> 32) & 1u);
        const long key = (long)(h.fkey & 0xffffffffull);
        if (kind == 0) {   // a boundary weight (src/rhs.jl:112-116)
            L.status = (key % 2 == 1) ? WEDG_ERR_BOUNDARY_TRACE : WEDG_ERR_BOUNDARY_DATA;
            L.key = key;
            L.side = (int)((key - 1) / (2 * nq * N));
            return L.status;
        }
        L.status = WEDG_ERR_MASS;   // src/rhs.jl:916
        L.element = key;
        return L.status;
    }
    if (h.nonfinite != INT_MAX) {
        L.step = h.nonfinite;
        if (nonfinite_is_error) L.status = WEDG_ERR_NONFINITE;
    }
    return L.status;
}

void sym_eigen(std::vector<double> A, int n, std::vector<double>& lam, std::vector<double>& V);
void stiffness_spectrum(const double* S, const double* w, int nu, std::vector<double>& Q,
                        std::vector<double>& lam);
bool all_equal(const double* p, size_t n);

}  // namespace wedg

// the public handles are global (include/wedg.h); their members use the kernels' types
using namespace wedg;

struct wedg_problem {
    int N = 0, E = 0, nu = 0, nv = 0, nh = 0, qu = 0;
    double omega = 0, theta = 0, tau = 0, lu = 0, lv = 0, mass_tol = 1e-14;
    int mass_cg = 0;   // conjugate gradients (mass_solver = :cg) instead of Chebyshev iteration
    int mms_on = 0, src_on = 0, one_sided = 0;
    double src_omega = 0, src_Ts = 0;
    wedg_side bc[4];
    int mode = MODE_GENERAL;   // kernel specialisation (wedg_kernels.h)
    int ncls_ca = 0, ncls_b = 0;
    std::vector<void*> allocs;
    double* tab = nullptr;
    double* tab2 = nullptr;   // split tables of the two-lane kernels (n_u = n_v even), or null
    int* cls_ca = nullptr;
    double *Cw = nullptr, *Aw = nullptr, *Ainv = nullptr;
    int* cls_b = nullptr;
    double* beta = nullptr;
    std::vector<double> lam, Q;   // spectrum of the 1D stiffness
    FaceArr f[F_COUNT];
    double fu[F_COUNT] = {};   // values of the face arrays that are one constant
    double* bcd[8] = {};
    double* proj[5] = {};   // PiSP, PiBSP2, PiCdiv, PiDdiv, src
    int has_proj[5] = {};
    size_t device_bytes = 0;
    // energy diagnostics (wedg_problem_set_diag): classes of (c^2, beta) at the standard points
    int* dcls = nullptr;
    double *dc2 = nullptr, *dbq = nullptr, *dW2d = nullptr;

    // Host data reaches the device only through this pinned buffer: copies
    // straight from pageable memory would make the runtime register (pin) the
    // caller's memory on the fly, which is not available on every system.
    void* stage = nullptr;
    size_t stage_bytes = 0;

    template <class T>
    T* upload(const T* h, size_t n, hipError_t* err) {
        T* d = dalloc<T>(n, err);
        if (*err != hipSuccess) return nullptr;
        allocs.push_back(d);
        device_bytes += n * sizeof(T);
        size_t done = 0, total = n * sizeof(T);
        while (done < total && *err == hipSuccess) {
            const size_t chunk = std::min(stage_bytes, total - done);
            std::memcpy(stage, reinterpret_cast<const char*>(h) + done, chunk);
            *err = hipMemcpy(reinterpret_cast<char*>(d) + done, stage, chunk, hipMemcpyHostToDevice);
            done += chunk;
        }
        return d;
    }
};

struct wedg_workspace {
    wedg_problem* p = nullptr;
    int cheb_global = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    double *yu = nullptr, *yv = nullptr, *su = nullptr, *sv = nullptr;
    double *au = nullptr, *av = nullptr, *ou = nullptr, *ov = nullptr;
    double* bsl = nullptr;   // boundary slots [side][i][A, D, F][N]
    DevErr* err = nullptr;
    DevErr* herr = nullptr;  // pinned host copy of the failure record
    double *hu = nullptr, *hv = nullptr;   // pinned staging of the state (host <-> device)
    int* cnt = nullptr;                 // per-wave Chebyshev counts [2][waves]
    unsigned* wstat = nullptr;          // per-wave count statistics [sum | max][waves]
    unsigned* hstat = nullptr;          // pinned host copy of wstat
    // source-mode diagnostics: device outputs and their pinned host copies
    double* drecv = nullptr; double* hrecv = nullptr; size_t nrecv_cap = 0;
    double* ddiag = nullptr; double* hdiag = nullptr;   // E, vmax, ndmin [3][E], slice [N nq], eta0 [nv]
    int* drecv_elem = nullptr; double* drecv_rows = nullptr; int nrecv_dev = 0;
    int nrecv_set = 0;       // receivers a batched advance records (wedg_set_receivers)
    void* hrows = nullptr;   // pinned staging of the receiver rows and elements
    double* dnorm = nullptr; double* hnorm = nullptr;   // error norms: [4][E] sums, then the 1D tables
    int serial = 0;
    long step_base = 0;
    long stages = 0;                 // stages run since the last reset
    unsigned long long cheb_wave_sum = 0;
    wedg_error_info last{};
    // timing
    int timing = 0;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_used;
    double t_ms[4] = {0, 0, 0, 0};
    long t_n[4] = {0, 0, 0, 0};
    // a HIP graph of the stage launches of graph_steps RK4 steps (wedg_graph_enable):
    // the captured graph, its kernel nodes in launch order with their launch
    // configuration, and a ring of executable copies whose arguments are set
    // before each launch (a copy is reused once its previous launch has ended)
    static constexpr int kGraphRing = 3;
    long graph_steps = 0;
    hipGraph_t graph = nullptr;
    const wedg_problem* graph_p = nullptr;
    std::vector<hipGraphNode_t> gnodes;
    std::vector<hipKernelNodeParams> gparams;
    hipGraphExec_t gexec[kGraphRing] = {};
    hipEvent_t gdone[kGraphRing] = {};
    bool gused[kGraphRing] = {};
    long glaunch = 0;
    // lanes per element of the stage kernels: 1 (k_stage) or 2 (k_stage2, where
    // the problem allows it; wedg_workspace_set_lanes_per_element)
    int lpe = 1;
    bool two() const {
        return lpe == 2 && p->tab2 && !p->mass_cg && (p->mode == MODE_UNIFORM || p->mode == MODE_SEPARABLE);
    }
    size_t nU() const { return (size_t)p->E * p->nu * p->nu; }
    // waves of the stage kernels this workspace runs (per-wave count slots and statistics)
    size_t nwaves() const { return ((size_t)p->E + (two() ? kPairWave : kWave) - 1) / (two() ? kPairWave : kWave); }
    size_t nwaves_cap() const { return ((size_t)p->E + kPairWave - 1) / kPairWave; }
    size_t nV() const { return (size_t)p->E * p->nv * p->nv; }
};

namespace wedg {

hipError_t launch(const KArgs& a, wedg_problem* p, int which, hipStream_t s, bool two);   // two: the two-lane kernels

// The host tables the split tables of the two-lane kernels are made from:
// column-major 1D tables as precompute forms them (w: nq, W2d_h: nh x nh,
// P_u/S_u: nq x nu, P_h_v: nh x nv, E rows: nu) and, for a one-class problem,
// its records (Cw, Aw: nq x nq [iy + nq ix]; beta: nh x nh [p + nh q]) and
// u-solve factors d0 (nu x nu); null where the problem has classes.
struct SplitSrc {
    const double *w, *W2d_h, *P_u, *S_u, *P_h_v, *Ebl, *Ebr, *E1bl, *E1br;
    const double *cw, *aw, *beta, *d0;
};
// Uploads p->tab2 when the two-lane kernels admit the problem (p->mode, p->Q
// and the sizes already set); leaves it null otherwise.
hipError_t build_split_tables(wedg_problem* p, const SplitSrc& src);
wedg_status reset_err(wedg_workspace* ws);
wedg_status collect(wedg_workspace* ws, int nonfinite_is_error);
KArgs problem_args(const wedg_problem* p, int cheb_global, double* bsl, DevErr* err, int* cnt, unsigned* wstat);
KArgs base_args(wedg_workspace* ws);
void stage_time(const wedg_problem* p, double t, KArgs& a);
wedg_status run_kernel(wedg_workspace* ws, const KArgs& a, int which);
KArgs stage_args(wedg_workspace* ws, int stage, const double* yu, const double* yv, const double* y0u,
                 const double* y0v, double* outu, double* outv, double t, double c, long step, bool has_next,
                 double t_next);
wedg_status rk4_device(wedg_workspace* ws, double t, double dt, long step, bool first, bool last);
wedg_status rk4_step_io(wedg_workspace* ws, const double* inu, const double* inv, double* outu, double* outv,
                        double t, double dt, long step, bool first, bool last);

}  // namespace wedg
