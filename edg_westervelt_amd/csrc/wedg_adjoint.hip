// Host side of the discrete adjoint (wedg_adj_* in include/wedg.h): source-mode
// runs with per-element coefficients c², α, β, the boundary misfit of u_h, its
// gradient with respect to c² and β by the adjoint of RK4 with binomial
// checkpointing, and single tangent-linear / vector-Jacobian products of the
// right-hand side.
//
// The forward steps run the forward stage kernels (k_prep, k_stage) on a
// workspace of a problem built here from the per-element arrays: the same
// tables precompute makes for coefficients that are constant on every element
// (src/solver.jl:509-552, 716-770), with one-sided face values where
// neighbours differ.
//
// The gradient runs three streams, each with its own workspace: the forward
// stream does the checkpointing schedule (advances, snapshots, restores) and
// copies the state of each step to be reversed into one of its tapes; the
// tape stream re-runs that step keeping its stage states and stage
// derivatives; the adjoint stream seeds the misfit of the step's end state and
// runs the eight adjoint kernels of the step. Events order a tape between its
// writers and its reader, so the forward stream runs ahead by as many steps to
// be reversed as there are tapes and the three chains overlap on the GPU (a
// 64 x 64 mesh fills a sixteenth of its SIMDs). Binomial checkpointing
// alternates stretches of many recomputed forward steps per reversal with
// stretches of few; the tapes let the adjoint chain work off in the first kind
// what it falls behind by in the second, so a gradient takes more tapes when
// they are small (up to kMaxTapes within kTapeBudget bytes, at least kMinTapes).
#include "wedg_internal.h"
#include "wedg_adjoint.h"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>
#include <math.h>

namespace {

hipError_t launch_adj2(const wedg::AdjArgs& b, const wedg::AdjArgs& a, int qu, int nv, hipStream_t s) {
    switch (qu) {
        case 1: return wedg::launch_adj2_q1(b, a, nv, s);
        case 2: return wedg::launch_adj2_q2(b, a, nv, s);
        case 3: return wedg::launch_adj2_q3(b, a, nv, s);
        case 4: return wedg::launch_adj2_q4(b, a, nv, s);
        case 5: return wedg::launch_adj2_q5(b, a, nv, s);
        case 6: return wedg::launch_adj2_q6(b, a, nv, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_adj(const wedg::AdjArgs& a, int qu, int nv, int which, hipStream_t s) {
    switch (qu) {
        case 1: return wedg::launch_adj_q1(a, nv, which, s);
        case 2: return wedg::launch_adj_q2(a, nv, which, s);
        case 3: return wedg::launch_adj_q3(a, nv, which, s);
        case 4: return wedg::launch_adj_q4(a, nv, which, s);
        case 5: return wedg::launch_adj_q5(a, nv, which, s);
        case 6: return wedg::launch_adj_q6(a, nv, which, s);
        default: return hipErrorInvalidValue;
    }
}

// binomial coefficient as a double (large arguments saturate harmlessly)
double binom(int n, int k) {
    if (k < 0 || k > n) return 0.0;
    double r = 1.0;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

// Optimal first split of n >= 2 steps reversed with s >= 1 free snapshot slots
// besides the one holding the segment's first state (binomial checkpointing,
// Griewank and Walther, ACM TOMS 26(1), 2000): with c = s + 1 snapshots and r
// the least repetition number with C(c + r, r) >= n, the left part of m steps
// is reversed with c snapshots and r - 1 repetitions, the right part with
// c - 1 and r, and the total number of forward steps, r n - C(c + r, r - 1), is
// reached for every m with C(c+r-2, r-2) <= m <= C(c+r-1, r-1) and
// C(c+r-2, r-1) <= n - m <= C(c+r-1, r).
long revolve_split(long n, int s) {
    const int c = s + 1;
    int r = 0;
    while (binom(c + r, r) < (double)n) ++r;
    double lo = std::max(1.0, std::max(binom(c + r - 2, r - 2), (double)n - binom(c + r - 1, r)));
    double hi = std::min((double)(n - 1), std::min(binom(c + r - 1, r - 1), (double)n - binom(c + r - 2, r - 1)));
    long m = (long)lo;
    if (!(lo <= hi)) m = std::max(1L, std::min(n - 1, (long)hi));
    return m;
}

// the state (u, v) copied in one launch
__global__ void k_copy_state(const double* __restrict__ su, double* __restrict__ du, size_t nu,
                             const double* __restrict__ sv, double* __restrict__ dv, size_t nv) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nu + nv; i += stride) {
        if (i < nu) du[i] = su[i];
        else dv[i - nu] = sv[i - nu];
    }
}

// States copied for several instances in one launch: job blockIdx.y copies
// (su, sv) into (du, dv).
struct CopyJob {
    const double* su;
    double* du;
    const double* sv;
    double* dv;
    size_t nu, nv;
};

__global__ void k_copy_states(const CopyJob* __restrict__ jobs) {
    const CopyJob J = jobs[blockIdx.y];
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < J.nu + J.nv; i += stride) {
        if (i < J.nu) J.du[i] = J.su[i];
        else J.dv[i - J.nu] = J.sv[i - J.nu];
    }
}

hipError_t copy_state(const double* su, double* du, size_t nu, const double* sv, double* dv, size_t nv,
                      hipStream_t s) {
    const size_t n = nu + nv;
    const int grid = (int)std::min<size_t>(1024, (n + 255) / 256);
    k_copy_state<<<grid, 256, 0, s>>>(su, du, nu, sv, dv, nv);
    return hipGetLastError();
}

}  // namespace

// tapes in flight between the forward and the adjoint stream: as many as fit
// in kTapeBudget bytes, between kMinTapes and kMaxTapes (a tape holds nine
// states: Y1..Y4, F(Y1)..F(Y4) and the end state)
constexpr int kMinTapes = 4, kMaxTapes = 32;
constexpr size_t kTapeBudget = (size_t)2 << 30;

struct AdjTape {
    double *Yu[4] = {}, *Yv[4] = {}, *Ku[4] = {}, *Kv[4] = {};   // stage states (Y1 a copy of y_n) and F(Y_s)
    double *ynu = nullptr, *ynv = nullptr;                        // the step's end state
    hipEvent_t copied = nullptr, taped = nullptr, used = nullptr; // y_n copied (forward), taped, reversed
};

struct wedg_adj {
    wedg_problem* p = nullptr;
    wedg_workspace* ws = nullptr;
    int N = 0, E = 0, nu = 0, nv = 0, nh = 0, qu = 0, cheb_policy = 0;
    double theta = 0, tau = 0, lu = 0, lv = 0, mass_tol = 1e-14, src_omega = 0, src_Ts = 0;
    wedg_side bc[4];
    std::vector<double> w, Pu, Su, Ebl, Ebr, E1bl, E1br, Ph, W2dh, src;   // host copies of the tables
    std::vector<double> alpha, c2, beta;                                     // [E], e = j + N k
    std::vector<double> lam, Q, K1;                                          // 1D stiffness
    std::vector<void*> allocs;
    double *dc2 = nullptr, *dal = nullptr, *dbe = nullptr, *atab = nullptr;
    std::vector<AdjTape> tape;   // tape[0] from creation (also the single products' scratch); more per gradient
    double *ybu = nullptr, *ybv = nullptr, *accu = nullptr, *accv = nullptr, *kbu = nullptr, *kbv = nullptr;
    double *pu = nullptr, *pv = nullptr, *slot = nullptr, *slot2 = nullptr, *gc2 = nullptr, *gbe = nullptr;
    double *tdc2 = nullptr, *tdbe = nullptr;   // tangent-linear directions of the coefficients
    int* fail = nullptr;
    double* snaps = nullptr; int nsnap_cap = 0;
    double* dtgt = nullptr; long tgt_steps = 0;
    double* djp = nullptr; double* dtr = nullptr; long jp_cap = 0;
    double* pin = nullptr; size_t pin_n = 0;   // pinned staging of host data
    wedg_workspace* wt = nullptr;              // the tape stream's workspace
    wedg_workspace* wa = nullptr;              // the adjoint stream's workspace
    hipStream_t sa = nullptr;                  // the adjoint stream (wa's)
    wedg_adj_stats stats{};
    size_t nU() const { return (size_t)E * nu * nu; }
    size_t nV() const { return (size_t)E * nv * nv; }
};

namespace {

// host <-> device through the pinned staging buffer, on the workspace stream
wedg_status h2d(wedg_adj* A, double* d, const double* h, size_t n) {
    hipStream_t s = A->ws->stream;
    for (size_t off = 0; off < n; off += A->pin_n) {
        const size_t k = std::min(A->pin_n, n - off);
        HIPOK(hipStreamSynchronize(s));
        std::memcpy(A->pin, h + off, k * sizeof(double));
        HIPOK(hipMemcpyAsync(d + off, A->pin, k * sizeof(double), hipMemcpyHostToDevice, s));
    }
    HIPOK(hipStreamSynchronize(s));
    return WEDG_OK;
}

wedg_status d2h(wedg_adj* A, double* h, const double* d, size_t n) {
    hipStream_t s = A->ws->stream;
    for (size_t off = 0; off < n; off += A->pin_n) {
        const size_t k = std::min(A->pin_n, n - off);
        HIPOK(hipMemcpyAsync(A->pin, d + off, k * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPOK(hipStreamSynchronize(s));
        std::memcpy(h + off, A->pin, k * sizeof(double));
    }
    return WEDG_OK;
}

double* adj_alloc(wedg_adj* A, size_t n, hipError_t* he) {
    if (*he != hipSuccess) return nullptr;
    double* d = dalloc<double>(n, he);
    if (*he == hipSuccess) A->allocs.push_back(d);
    return d;
}

// Tapes up to n (the ones there are kept): device buffers and their three events.
wedg_status ensure_tapes(wedg_adj* A, int n) {
    const size_t nU = A->nU(), nV = A->nV();
    while ((int)A->tape.size() < n) {
        A->tape.emplace_back();
        AdjTape& tp = A->tape.back();
        hipError_t he = hipSuccess;
        for (int s = 0; s < 4; ++s) {
            tp.Yu[s] = adj_alloc(A, nU, &he); tp.Yv[s] = adj_alloc(A, nV, &he);
            tp.Ku[s] = adj_alloc(A, nU, &he); tp.Kv[s] = adj_alloc(A, nV, &he);
        }
        tp.ynu = adj_alloc(A, nU, &he); tp.ynv = adj_alloc(A, nV, &he);
        if (he == hipSuccess) he = hipEventCreateWithFlags(&tp.copied, hipEventDisableTiming);
        if (he == hipSuccess) he = hipEventCreateWithFlags(&tp.taped, hipEventDisableTiming);
        if (he == hipSuccess) he = hipEventCreateWithFlags(&tp.used, hipEventDisableTiming);
        if (he != hipSuccess) {
            // a tape that is not whole is not kept (its buffers stay in allocs until destroy)
            for (hipEvent_t ev : {tp.copied, tp.taped, tp.used})
                if (ev) (void)hipEventDestroy(ev);
            A->tape.pop_back();
            return WEDG_ERR_HIP;
        }
    }
    return WEDG_OK;
}

// The tapes of a gradient of nsteps steps (at most one per step).
int tapes_for(const wedg_adj* A, long nsteps) {
    const size_t bytes = (size_t)9 * (A->nU() + A->nV()) * sizeof(double);
    const long fit = (long)(kTapeBudget / std::max<size_t>(bytes, 1));
    const long n = std::min<long>(kMaxTapes, std::max<long>(kMinTapes, fit));
    return (int)std::max(1L, std::min(n, nsteps));
}

// The forward problem of the current coefficients: the tables of precompute for
// coefficients constant on every element (src/solver.jl:509-552, 647-662, 716-770),
// built from the per-element arrays directly.
wedg_status build_problem(wedg_adj* A, wedg_problem** out) {
    *out = nullptr;
    const int N = A->N, nu = A->nu, nv = A->nv, nh = A->nh, nq = nu;
    const long E = A->E;
    auto* p = new wedg_problem();
    p->stage_bytes = (size_t)4 << 20;
    if (hipHostMalloc(&p->stage, p->stage_bytes, hipHostMallocDefault) != hipSuccess) {
        delete p;
        return WEDG_ERR_HIP;
    }
    p->N = N; p->E = (int)E; p->nu = nu; p->nv = nv; p->nh = nh; p->qu = nu - 1;
    p->omega = 0.0; p->theta = A->theta; p->tau = A->tau; p->lu = A->lu; p->lv = A->lv;
    p->mass_tol = A->mass_tol; p->mass_cg = 0; p->mms_on = 0; p->src_on = 1;
    p->src_omega = A->src_omega; p->src_Ts = A->src_Ts;
    for (int s = 0; s < 4; ++s) p->bc[s] = A->bc[s];
    p->lam = A->lam;
    p->Q = A->Q;
    hipError_t he = hipSuccess;
#define UP(dst, src, n)                                                     \
    do {                                                                    \
        dst = p->upload(src, n, &he);                                       \
        if (he != hipSuccess) { wedg_problem_destroy(p); return WEDG_ERR_HIP; } \
    } while (0)
    {
        std::vector<double> t;
        t.insert(t.end(), A->Pu.begin(), A->Pu.end());
        t.insert(t.end(), A->Su.begin(), A->Su.end());
        t.insert(t.end(), A->w.begin(), A->w.end());
        t.insert(t.end(), A->Ebl.begin(), A->Ebl.end());
        t.insert(t.end(), A->Ebr.begin(), A->Ebr.end());
        t.insert(t.end(), A->E1bl.begin(), A->E1bl.end());
        t.insert(t.end(), A->E1br.begin(), A->E1br.end());
        t.insert(t.end(), A->Ph.begin(), A->Ph.end());
        t.insert(t.end(), A->W2dh.begin(), A->W2dh.end());
        t.insert(t.end(), A->Q.begin(), A->Q.end());
        const double h = 1.0 / N, two_h = 2.0 / h, four_hh = 4.0 / (h * h);
        for (int a = 0; a < nu; ++a) t.push_back(two_h * A->Ebl[a]);
        for (int a = 0; a < nu; ++a) t.push_back(two_h * A->Ebr[a]);
        for (int a = 0; a < nu; ++a) t.push_back(four_hh * A->E1bl[a]);
        for (int a = 0; a < nu; ++a) t.push_back(four_hh * A->E1br[a]);
        UP(p->tab, t.data(), t.size());
    }
    // classes of (c^2, alpha): Cw = c^2 w w^T, Aw = alpha w w^T and the spectral
    // factors 1 / (c^2 (lam_i + lam_j) + [i = j = 0]) of the pinned stiffness
    {
        std::vector<int> cls;
        std::vector<double> uniq;
        dedup(E, 2, [&](long e, double* b) { b[0] = A->c2[e]; b[1] = A->alpha[e]; }, cls, uniq);
        const int nc = (int)(uniq.size() / 2);
        std::vector<double> cw((size_t)nc * nq * nq), aw((size_t)nc * nq * nq), dfac((size_t)nc * nu * nu);
        for (int c = 0; c < nc; ++c) {
            const double c2 = uniq[2 * c], al = uniq[2 * c + 1];
            for (int ix = 0; ix < nq; ++ix)
                for (int iy = 0; iy < nq; ++iy) {
                    const double ww = A->w[iy] * A->w[ix];
                    cw[(size_t)c * nq * nq + iy + nq * ix] = c2 * ww;
                    aw[(size_t)c * nq * nq + iy + nq * ix] = al * ww;
                }
            for (int jj = 0; jj < nu; ++jj)
                for (int i = 0; i < nu; ++i)
                    dfac[(size_t)c * nu * nu + i + nu * jj] =
                        1.0 / (c2 * (A->lam[i] + A->lam[jj]) + (i == 0 && jj == 0 ? 1.0 : 0.0));
        }
        p->ncls_ca = nc;
        UP(p->Cw, cw.data(), cw.size());
        UP(p->Aw, aw.data(), aw.size());
        UP(p->Ainv, dfac.data(), dfac.size());
        UP(p->cls_ca, cls.data(), cls.size());
    }
    {
        std::vector<int> cls;
        std::vector<double> uniq;
        dedup(E, 1, [&](long e, double* b) { b[0] = A->beta[e]; }, cls, uniq);
        const int nc = (int)uniq.size();
        std::vector<double> bt((size_t)nc * nh * nh);
        for (int c = 0; c < nc; ++c)
            for (int i = 0; i < nh * nh; ++i) bt[(size_t)c * nh * nh + i] = uniq[c];
        p->ncls_b = nc;
        UP(p->beta, bt.data(), bt.size());
        UP(p->cls_b, cls.data(), cls.size());
    }
    // face arrays [i][face]: horizontal faces jf + (N + 1) k, vertical j + N kf; the
    // "+" element's values (the inside element's on the physical boundary) and the
    // "-" element's; c on the boundary sqrt(c^2) (src/solver.jl:546-552, 716-751)
    bool one_sided = false;
    {
        const size_t nf = (size_t)N * (N + 1);
        auto el = [&](const std::vector<double>& f, long j, long k) { return f[j + (long)N * k]; };
        std::vector<double> cfv(E);
        for (long e = 0; e < E; ++e) cfv[e] = std::sqrt(A->c2[e]);
        auto hplus = [&](const std::vector<double>& f, long jf, long k) { return el(f, jf < N ? jf : N - 1, k); };
        auto hminus = [&](const std::vector<double>& f, long jf, long k) { return el(f, jf > 0 ? jf - 1 : 0, k); };
        auto vplus = [&](const std::vector<double>& f, long j, long kf) { return el(f, j, kf < N ? kf : N - 1); };
        auto vminus = [&](const std::vector<double>& f, long j, long kf) { return el(f, j, kf > 0 ? kf - 1 : 0); };
        std::vector<double> arr[F_COUNT];
        for (int a = 0; a < F_COUNT; ++a) arr[a].assign(nf, 0.0);
        for (long k = 0; k < N; ++k)
            for (long jf = 0; jf <= N; ++jf) {
                const size_t f = jf + (size_t)(N + 1) * k;
                arr[F_HC2][f] = hplus(A->c2, jf, k);
                arr[F_HAF][f] = hplus(A->alpha, jf, k);
                arr[F_HCF][f] = hplus(cfv, jf, k);
                arr[F_HBF][f] = hplus(A->beta, jf, k);
                arr[F_HC2M][f] = hminus(A->c2, jf, k);
                arr[F_HAFM][f] = hminus(A->alpha, jf, k);
                if (jf > 0 && jf < N &&
                    (arr[F_HC2][f] != arr[F_HC2M][f] || arr[F_HAF][f] != arr[F_HAFM][f])) one_sided = true;
            }
        for (long kf = 0; kf <= N; ++kf)
            for (long j = 0; j < N; ++j) {
                const size_t f = j + (size_t)N * kf;
                arr[F_VC2][f] = vplus(A->c2, j, kf);
                arr[F_VAF][f] = vplus(A->alpha, j, kf);
                arr[F_VCF][f] = vplus(cfv, j, kf);
                arr[F_VBF][f] = vplus(A->beta, j, kf);
                arr[F_VC2M][f] = vminus(A->c2, j, kf);
                arr[F_VAFM][f] = vminus(A->alpha, j, kf);
                if (kf > 0 && kf < N &&
                    (arr[F_VC2][f] != arr[F_VC2M][f] || arr[F_VAF][f] != arr[F_VAFM][f])) one_sided = true;
            }
        bool faces_uni = !one_sided;
        for (int a = 0; a < F_COUNT; ++a) {
            if (all_equal(arr[a].data(), nf)) {
                double v = arr[a][0];
                UP(p->f[a].p, &v, 1);
                p->f[a].uni = 1;
                p->fu[a] = v;
            } else {
                std::vector<double> t(nf * nq);   // constant over the points of a face
                for (int i = 0; i < nq; ++i) std::memcpy(&t[(size_t)i * nf], arr[a].data(), nf * sizeof(double));
                UP(p->f[a].p, t.data(), t.size());
                p->f[a].uni = 0;
                faces_uni = false;
            }
        }
        p->one_sided = one_sided ? 1 : 0;
        p->mode = (p->ncls_ca == 1 && p->ncls_b == 1 && faces_uni) ? MODE_UNIFORM : MODE_SEPARABLE;
    }
    // the split tables of the two-lane stage kernels (the forward steps may run them)
    {
        SplitSrc src{A->w.data(), A->W2dh.data(), A->Pu.data(), A->Su.data(), A->Ph.data(), A->Ebl.data(),
                     A->Ebr.data(), A->E1bl.data(), A->E1br.data(), nullptr, nullptr, nullptr, nullptr};
        std::vector<double> rc((size_t)nq * nq), ra((size_t)nq * nq), rb((size_t)nh * nh, A->beta[0]),
            d0((size_t)nu * nu);
        if (p->mode == MODE_UNIFORM) {
            for (int ix = 0; ix < nq; ++ix)
                for (int iy = 0; iy < nq; ++iy) {
                    rc[iy + (size_t)nq * ix] = A->c2[0] * A->w[iy] * A->w[ix];
                    ra[iy + (size_t)nq * ix] = A->alpha[0] * A->w[iy] * A->w[ix];
                }
            for (int jj = 0; jj < nu; ++jj)
                for (int i = 0; i < nu; ++i)
                    d0[i + (size_t)nu * jj] = 1.0 / (A->c2[0] * (A->lam[i] + A->lam[jj]) + (i == 0 && jj == 0 ? 1.0 : 0.0));
            src.cw = rc.data(); src.aw = ra.data(); src.beta = rb.data(); src.d0 = d0.data();
        }
        if (build_split_tables(p, src) != hipSuccess) { wedg_problem_destroy(p); return WEDG_ERR_HIP; }
    }
    {
        std::vector<double> z((size_t)nq * N, 0.0);   // zero boundary data (src/solver.jl:647-655)
        for (int a = 0; a < 8; ++a) UP(p->bcd[a], z.data(), z.size());
    }
    {
        const size_t n2 = (size_t)nv * nv;
        std::vector<double> t(n2 * E);
        bool zero = true;
        for (long e = 0; e < E; ++e)
            for (size_t m = 0; m < n2; ++m) {
                const double v = A->src[m + n2 * e];
                t[m * E + e] = v;
                if (v != 0.0) zero = false;
            }
        if (!zero) {
            UP(p->proj[4], t.data(), t.size());
            p->has_proj[4] = 1;
        }
    }
#undef UP
    (void)hipHostFree(p->stage);
    p->stage = nullptr;
    *out = p;
    return WEDG_OK;
}

AdjArgs adj_args(wedg_adj* A) {
    AdjArgs a;
    std::memset(&a, 0, sizeof(a));
    a.N = A->N; a.E = A->E;
    const double h = 1.0 / A->N;
    a.two_h = 2.0 / h; a.four_hh = 4.0 / (h * h); a.hw = 0.5 * h;
    a.theta = A->theta; a.tau = A->tau; a.lu = A->lu; a.lv = A->lv;
    a.cheb_L = std::acosh(1.0 / A->mass_tol);
    for (int s = 0; s < 4; ++s) {
        a.kind[s] = A->bc[s].kind; a.gam[s] = A->bc[s].gamma_abc; a.eta[s] = A->bc[s].eta_abc;
        a.sig[s] = A->bc[s].sigma; a.kb[s] = A->bc[s].kb; a.kap[s] = A->bc[s].kappa;
    }
    a.tab = A->p->tab; a.atab = A->atab;
    a.c2 = A->dc2; a.al = A->dal; a.be = A->dbe;
    a.slot = A->slot; a.slotr = A->slot; a.pu = A->pu; a.pv = A->pv; a.gc2 = A->gc2; a.gbe = A->gbe;
    a.ybu = A->ybu; a.ybv = A->ybv; a.accu = A->accu; a.accv = A->accv; a.nkbu = A->kbu; a.nkbv = A->kbv;
    a.fail = A->fail;
    a.kscale = 1.0;
    return a;
}

// One RK4 step from the tape's Y1 = y_n (time t) on the stream of workspace ws,
// taped: the stage states Y2 .. Y4, the stage derivatives k1 .. k4 and the end
// state, as the forward steps compute them (src/solver.jl:289-321).
wedg_status tape_step(wedg_workspace* ws, AdjTape& tp, double t, double dt, long step) {
    const double hd = 0.5 * dt, th = t + 0.5 * dt, tf = t + dt;
    const double ts[4] = {t, th, th, tf}, cs[4] = {hd, hd, dt, dt / 6.0}, tn[4] = {th, th, tf, tf};
    wedg_status st;
    for (int s = 0; s < 4; ++s) {
        double* ou = s < 3 ? tp.Yu[s + 1] : tp.ynu;
        double* ov = s < 3 ? tp.Yv[s + 1] : tp.ynv;
        KArgs a = stage_args(ws, s + 1, tp.Yu[s], tp.Yv[s], tp.Yu[0], tp.Yv[0], ou, ov, ts[s], cs[s], step, s < 3,
                             tn[s]);
        a.recu = tp.Ku[s];
        a.recv = tp.Kv[s];
        if (s == 0 && (st = run_kernel(ws, a, 0)) != WEDG_OK) return st;
        if ((st = run_kernel(ws, a, 1)) != WEDG_OK) return st;
    }
    return WEDG_OK;
}

// The arguments of the adjoint of one taped step: the misfit of its end state
// (target row n) into ybar (mis), and stages 4..1 (st[1..4]); Y1, the step's
// start state, is (y1u, y1v). The slot adjoints alternate between two buffers.
void adjoint_step_args(wedg_adj* A, const AdjTape& tp, long n, double dt, const double* y1u, const double* y1v,
                       AdjArgs& mis, AdjArgs (&st)[5]) {
    AdjArgs a = adj_args(A);
    mis = a;
    mis.yu = tp.ynu;
    mis.tgt = A->dtgt ? A->dtgt + (size_t)n * A->N * A->nu : nullptr;
    mis.jp = A->dtgt ? A->djp + (size_t)n * A->N * A->nu : nullptr;
    mis.ybu = A->ybu;
    const double ca[5] = {0, 0, dt / 6.0, dt / 3.0, dt / 3.0};   // next cotangent: ca ybar + cb Ybar_s
    const double cb[5] = {0, 0, 0.5 * dt, 0.5 * dt, dt};
    for (int stg = 4; stg >= 1; --stg) {
        AdjArgs& x = st[stg];
        x = a;
        x.stage = stg;
        x.yu = stg == 1 ? y1u : tp.Yu[stg - 1];
        x.yv = stg == 1 ? y1v : tp.Yv[stg - 1];
        x.ku = tp.Ku[stg - 1]; x.kv = tp.Kv[stg - 1];
        if (stg == 4) { x.kbu = A->ybu; x.kbv = A->ybv; x.kscale = dt / 6.0; }
        else { x.kbu = A->kbu; x.kbv = A->kbv; x.kscale = 1.0; }
        x.ca = ca[stg]; x.cb = cb[stg];
        x.slot = (stg & 1) ? A->slot2 : A->slot;   // written by the stage's k_adj_a
        x.slotr = x.slot;                          // read by its k_adj_b
    }
    st[0] = a;
}

// The adjoint of one taped step on the adjoint stream: k_misfit, then k_adj_a
// of stage 4, three launches of k_adj_b of stage s with k_adj_a of stage s - 1,
// and k_adj_b of stage 1.
wedg_status adjoint_step(wedg_adj* A, AdjTape& tp, long n, double dt) {
    const hipStream_t s = A->sa;
    AdjArgs mis, st[5];
    adjoint_step_args(A, tp, n, dt, tp.Yu[0], tp.Yv[0], mis, st);
    if (A->dtgt) HIPOK(launch_adj(mis, A->qu, A->nv, 3, s));
    HIPOK(launch_adj(st[4], A->qu, A->nv, 0, s));
    for (int stg = 4; stg >= 2; --stg) HIPOK(launch_adj2(st[stg], st[stg - 1], A->qu, A->nv, s));
    HIPOK(launch_adj(st[1], A->qu, A->nv, 1, s));
    ++A->stats.adjoint_steps;
    return WEDG_OK;
}

wedg_status zero_dev(hipStream_t s, double* d, size_t n) {
    HIPOK(hipMemsetAsync(d, 0, n * sizeof(double), s));
    return WEDG_OK;
}

wedg_status read_fail(wedg_adj* A) {
    int f = 0;
    HIPOK(hipStreamSynchronize(A->sa));
    HIPOK(hipStreamSynchronize(A->ws->stream));
    HIPOK(hipMemcpy(&f, A->fail, sizeof(int), hipMemcpyDeviceToHost));
    return f ? WEDG_ERR_MASS : WEDG_OK;
}

wedg_status jp_capacity(wedg_adj* A, long nsteps) {
    const long need = nsteps * A->N * A->nu;
    if (need <= A->jp_cap) return WEDG_OK;
    HIPOK(hipStreamSynchronize(A->ws->stream));
    HIPOK(hipStreamSynchronize(A->sa));
    if (A->djp) (void)hipFree(A->djp);
    if (A->dtr) (void)hipFree(A->dtr);
    A->djp = nullptr; A->dtr = nullptr; A->jp_cap = 0;
    HIPOK(hipMalloc((void**)&A->djp, need * sizeof(double)));
    HIPOK(hipMalloc((void**)&A->dtr, need * sizeof(double)));
    A->jp_cap = need;
    return WEDG_OK;
}

// J = sum over steps and points of the per-point terms, in step order
wedg_status sum_misfit(wedg_adj* A, long nsteps, double* J) {
    const size_t n = (size_t)nsteps * A->N * A->nu;
    std::vector<double> h(n);
    wedg_status s;
    if ((s = d2h(A, h.data(), A->djp, n))) return s;
    double acc = 0.0;
    for (size_t i = 0; i < n; ++i) acc += h[i];
    *J = acc;
    return WEDG_OK;
}

}  // namespace

extern "C" wedg_status wedg_adj_create(const wedg_adj_desc* d, int cheb_policy, wedg_adj** out) {
    if (!d || !out) return WEDG_ERR_ARGUMENT;
    *out = nullptr;
    const int N = d->N, nu = d->qu1, nv = d->qv1;
    if (N < 2 || nu < 2 || nu > 7 || (nv != nu && nv != nu - 1)) return WEDG_ERR_ARGUMENT;
    if (!d->w || !d->P_u || !d->S_u || !d->Ebl_u || !d->Ebr_u || !d->E1bl_u || !d->E1br_u || !d->P_h_v ||
        !d->W2d_h || !d->src_projE || !d->alpha)
        return WEDG_ERR_ARGUMENT;
    if (cheb_policy != WEDG_CHEB_GLOBAL && cheb_policy != WEDG_CHEB_ELEMENT) return WEDG_ERR_ARGUMENT;
    const int qv = nv - 1;
    int nh = (3 * qv + 1) / 2 + 1;
    if (nh < nu) nh = nu;
    auto* A = new wedg_adj();
    A->N = N; A->E = N * N; A->nu = nu; A->nv = nv; A->nh = nh; A->qu = nu - 1; A->cheb_policy = cheb_policy;
    A->theta = d->theta; A->tau = d->tau; A->lu = d->lambda_u; A->lv = d->lambda_v;
    A->mass_tol = d->mass_tol > 0 ? d->mass_tol : 1e-14;
    A->src_omega = d->src_omega; A->src_Ts = d->src_Ts;
    for (int s = 0; s < 4; ++s) A->bc[s] = d->bc[s];
    const int nq = nu;
    auto cp = [](const double* p, size_t n) { return std::vector<double>(p, p + n); };
    A->w = cp(d->w, nq); A->Pu = cp(d->P_u, (size_t)nq * nu); A->Su = cp(d->S_u, (size_t)nq * nu);
    A->Ebl = cp(d->Ebl_u, nu); A->Ebr = cp(d->Ebr_u, nu); A->E1bl = cp(d->E1bl_u, nu); A->E1br = cp(d->E1br_u, nu);
    A->Ph = cp(d->P_h_v, (size_t)nh * nv); A->W2dh = cp(d->W2d_h, (size_t)nh * nh);
    A->src = cp(d->src_projE, (size_t)nv * nv * A->E);
    A->alpha = cp(d->alpha, A->E);
    A->c2.assign(A->E, 1.0);
    A->beta.assign(A->E, 0.0);
    stiffness_spectrum(d->S_u, d->w, nu, A->Q, A->lam);
    A->K1.assign((size_t)nu * nu, 0.0);
    for (int a = 0; a < nu; ++a)
        for (int b = 0; b < nu; ++b) {
            double s = 0.0;
            for (int p = 0; p < nq; ++p) s += d->w[p] * d->S_u[p + nq * a] * d->S_u[p + nq * b];
            A->K1[a + (size_t)nu * b] = s;
        }
    auto fail = [&](wedg_status s) { wedg_adj_destroy(A); return s; };
    wedg_status st;
    if ((st = build_problem(A, &A->p))) return fail(st);
    if ((st = wedg_workspace_create(A->p, cheb_policy, &A->ws))) return fail(st);
    if ((st = wedg_workspace_create(A->p, cheb_policy, &A->wt))) return fail(st);
    if ((st = wedg_workspace_create(A->p, cheb_policy, &A->wa))) return fail(st);
    A->sa = A->wa->stream;
    hipError_t he = hipSuccess;
    A->pin_n = std::max<size_t>((size_t)1 << 18, (size_t)A->E);
    if (hipHostMalloc((void**)&A->pin, A->pin_n * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        A->pin = nullptr;
        return fail(WEDG_ERR_HIP);
    }
    const size_t E = A->E, nU = A->nU(), nV = A->nV();
    A->dc2 = adj_alloc(A, E, &he); A->dal = adj_alloc(A, E, &he); A->dbe = adj_alloc(A, E, &he);
    A->tdc2 = adj_alloc(A, E, &he); A->tdbe = adj_alloc(A, E, &he);
    A->atab = adj_alloc(A, (size_t)nq * nq + (size_t)nu * nu + nu, &he);
    if (ensure_tapes(A, 1) != WEDG_OK) return fail(WEDG_ERR_HIP);
    A->ybu = adj_alloc(A, nU, &he); A->ybv = adj_alloc(A, nV, &he);
    A->accu = adj_alloc(A, nU, &he); A->accv = adj_alloc(A, nV, &he);
    A->kbu = adj_alloc(A, nU, &he); A->kbv = adj_alloc(A, nV, &he);
    A->pu = adj_alloc(A, nU, &he); A->pv = adj_alloc(A, nV, &he);
    A->slot = adj_alloc(A, (size_t)12 * nq * E, &he);
    A->slot2 = adj_alloc(A, (size_t)12 * nq * E, &he);
    A->gc2 = adj_alloc(A, E, &he); A->gbe = adj_alloc(A, E, &he);
    if (he == hipSuccess) {
        void* f = nullptr;
        he = hipMalloc(&f, 16);
        if (he == hipSuccess) { A->allocs.push_back(f); A->fail = (int*)f; }
    }
    if (he != hipSuccess) return fail(WEDG_ERR_HIP);
    std::vector<double> at;
    for (int q = 0; q < nq; ++q)
        for (int p = 0; p < nq; ++p) at.push_back(A->w[p] * A->w[q]);
    at.insert(at.end(), A->K1.begin(), A->K1.end());
    at.insert(at.end(), A->lam.begin(), A->lam.end());
    if ((st = h2d(A, A->atab, at.data(), at.size()))) return fail(st);
    if ((st = h2d(A, A->dal, A->alpha.data(), E))) return fail(st);
    if ((st = wedg_adj_set_params(A, A->c2.data(), A->beta.data()))) return fail(st);
    *out = A;
    return WEDG_OK;
}

extern "C" wedg_status wedg_adj_set_lanes_per_element(wedg_adj* A, int lanes) {
    if (!A || (lanes != 1 && lanes != 2)) return WEDG_ERR_ARGUMENT;
    // every workspace that runs forward stages: the forward, tape and adjoint streams'
    for (wedg_workspace* w : {A->ws, A->wt, A->wa}) {
        const wedg_status s = wedg_workspace_set_lanes_per_element(w, lanes);
        if (s) return s;
    }
    return WEDG_OK;
}

extern "C" wedg_status wedg_adj_lanes_per_element(const wedg_adj* A, int* lanes) {
    if (!A || !lanes) return WEDG_ERR_ARGUMENT;
    return wedg_workspace_lanes_per_element(A->ws, lanes);
}

extern "C" wedg_status wedg_adj_destroy(wedg_adj* A) {
    if (!A) return WEDG_OK;
    if (A->sa) (void)hipStreamSynchronize(A->sa);
    if (A->ws) (void)hipStreamSynchronize(A->ws->stream);
    for (auto& tp : A->tape) {
        if (tp.copied) (void)hipEventDestroy(tp.copied);
        if (tp.taped) (void)hipEventDestroy(tp.taped);
        if (tp.used) (void)hipEventDestroy(tp.used);
    }
    for (void* a : A->allocs) (void)hipFree(a);
    if (A->snaps) (void)hipFree(A->snaps);
    if (A->dtgt) (void)hipFree(A->dtgt);
    if (A->djp) (void)hipFree(A->djp);
    if (A->dtr) (void)hipFree(A->dtr);
    if (A->pin) (void)hipHostFree(A->pin);
    if (A->wt) (void)hipStreamSynchronize(A->wt->stream);
    if (A->wa) wedg_workspace_destroy(A->wa);
    if (A->wt) wedg_workspace_destroy(A->wt);
    if (A->ws) wedg_workspace_destroy(A->ws);
    if (A->p) wedg_problem_destroy(A->p);
    delete A;
    return WEDG_OK;
}

extern "C" wedg_status wedg_adj_set_params(wedg_adj* A, const double* c2, const double* beta) {
    if (!A || !c2 || !beta) return WEDG_ERR_ARGUMENT;
    for (int e = 0; e < A->E; ++e)
        if (!(c2[e] > 0.0) || !std::isfinite(c2[e]) || !std::isfinite(beta[e])) return WEDG_ERR_ARGUMENT;
    if (A->ws) {
        HIPOK(hipStreamSynchronize(A->ws->stream));
        HIPOK(hipStreamSynchronize(A->sa));
    }
    A->c2.assign(c2, c2 + A->E);
    A->beta.assign(beta, beta + A->E);
    wedg_problem* np = nullptr;
    wedg_status st = build_problem(A, &np);
    if (st) return st;
    if (A->ws) A->ws->p = np;
    if (A->wa) A->wa->p = np;
    if (A->wt) A->wt->p = np;
    if (A->p) wedg_problem_destroy(A->p);
    A->p = np;
    if ((st = h2d(A, A->dc2, A->c2.data(), A->E))) return st;
    return h2d(A, A->dbe, A->beta.data(), A->E);
}

extern "C" wedg_status wedg_adj_set_target(wedg_adj* A, long nsteps, const double* target) {
    if (!A || nsteps < 0 || (nsteps > 0 && !target)) return WEDG_ERR_ARGUMENT;
    HIPOK(hipStreamSynchronize(A->ws->stream));
    HIPOK(hipStreamSynchronize(A->sa));
    if (A->dtgt) (void)hipFree(A->dtgt);
    A->dtgt = nullptr;
    A->tgt_steps = 0;
    if (nsteps == 0) return WEDG_OK;
    const size_t n = (size_t)nsteps * A->N * A->nu;
    HIPOK(hipMalloc((void**)&A->dtgt, n * sizeof(double)));
    A->tgt_steps = nsteps;
    return h2d(A, A->dtgt, target, n);
}

extern "C" wedg_status wedg_adj_forward(wedg_adj* A, double dt, long nsteps, double* J, double* traces) {
    if (!A || nsteps < 1 || !(dt > 0)) return WEDG_ERR_ARGUMENT;
    if (J && A->tgt_steps < nsteps) return WEDG_ERR_ARGUMENT;
    wedg_workspace* ws = A->ws;
    wedg_status st;
    if ((st = jp_capacity(A, nsteps))) return st;
    if ((st = reset_err(ws))) return st;
    HIPOK(hipMemsetAsync(ws->yu, 0, A->nU() * sizeof(double), ws->stream));
    HIPOK(hipMemsetAsync(ws->yv, 0, A->nV() * sizeof(double), ws->stream));
    AdjArgs a = adj_args(A);
    a.yu = ws->yu;
    double t = 0.0;
    const size_t row = (size_t)A->N * A->nu;
    for (long n = 0; n < nsteps; ++n) {
        if ((st = rk4_device(ws, t, dt, n + 1, n == 0, n == nsteps - 1))) return st;
        t += dt;
        a.tgt = J ? A->dtgt + n * row : nullptr;
        a.jp = J ? A->djp + n * row : nullptr;
        a.tr = traces ? A->dtr + n * row : nullptr;
        HIPOK(launch_adj(a, A->qu, A->nv, 3, ws->stream));
    }
    if ((st = collect(ws, 1))) return st;
    if (J && (st = sum_misfit(A, nsteps, J))) return st;
    if (traces && (st = d2h(A, traces, A->dtr, nsteps * row))) return st;
    return WEDG_OK;
}

extern "C" wedg_status wedg_adj_gradient(wedg_adj* A, double dt, long nsteps, int nsnap, double* J,
                                         double* dJ_dc2, double* dJ_dbeta) {
    if (!A || nsteps < 1 || !(dt > 0) || nsnap < 1 || !J || !dJ_dc2 || !dJ_dbeta) return WEDG_ERR_ARGUMENT;
    if (A->tgt_steps < nsteps) return WEDG_ERR_ARGUMENT;
    wedg_workspace* ws = A->ws;
    const hipStream_t fs = ws->stream, as = A->sa;
    wedg_status st;
    if ((st = jp_capacity(A, nsteps))) return st;
    const size_t nU = A->nU(), nV = A->nV(), sz = nU + nV;
    if (nsnap > A->nsnap_cap) {
        HIPOK(hipStreamSynchronize(fs));
        if (A->snaps) (void)hipFree(A->snaps);
        A->snaps = nullptr;
        A->nsnap_cap = 0;
        HIPOK(hipMalloc((void**)&A->snaps, (size_t)nsnap * sz * sizeof(double)));
        A->nsnap_cap = nsnap;
    }
    std::memset(&A->stats, 0, sizeof(A->stats));
    const int ntapes = tapes_for(A, nsteps);
    if ((st = ensure_tapes(A, ntapes))) return st;
    A->stats.tapes = ntapes;
    if ((st = reset_err(ws)) || (st = reset_err(A->wt)) || (st = reset_err(A->wa))) return st;
    HIPOK(hipMemsetAsync(A->fail, 0, sizeof(int), fs));
    HIPOK(hipMemsetAsync(ws->yu, 0, nU * sizeof(double), fs));
    HIPOK(hipMemsetAsync(ws->yv, 0, nV * sizeof(double), fs));
    if ((st = zero_dev(fs, A->gc2, A->E)) || (st = zero_dev(fs, A->gbe, A->E)) || (st = zero_dev(fs, A->ybu, nU)) ||
        (st = zero_dev(fs, A->ybv, nV)))
        return st;
    HIPOK(hipStreamSynchronize(fs));
    // times after n steps, accumulated as the forward loop does (src/solver.jl:213-216)
    std::vector<double> tt(nsteps + 1);
    tt[0] = 0.0;
    for (long n = 0; n < nsteps; ++n) tt[n + 1] = tt[n] + dt;
    // The current state lives in a snapshot slot (after a restore) or in the
    // workspace state; an advance reads it where it is and writes its last step
    // into the workspace state or straight into the snapshot slot it is saved in.
    auto snap_u = [&](int slot) { return A->snaps + (size_t)slot * sz; };
    auto snap_v = [&](int slot) { return A->snaps + (size_t)slot * sz + nU; };
    const double* cur_u = snap_u(0);
    const double* cur_v = snap_v(0);
    auto restore = [&](int slot) { cur_u = snap_u(slot); cur_v = snap_v(slot); };
    auto advance = [&](long a, long m, int save_slot) -> wedg_status {
        wedg_status s2;
        for (long i = a; i < m; ++i) {
            const bool into_slot = i == m - 1 && save_slot >= 0;
            double* ou = into_slot ? snap_u(save_slot) : ws->yu;
            double* ov = into_slot ? snap_v(save_slot) : ws->yv;
            if ((s2 = rk4_step_io(ws, cur_u, cur_v, ou, ov, tt[i], dt, i + 1, i == a, i == m - 1))) return s2;
            cur_u = ou;
            cur_v = ov;
            ++A->stats.forward_steps;
        }
        if (save_slot >= 0) ++A->stats.snapshots_taken;
        return WEDG_OK;
    };
    // reverse step n (the current state is y_n): the forward stream copies y_n into a
    // free tape, the tape stream re-runs the step keeping its stages, the adjoint
    // stream reverses it
    const hipStream_t ts = A->wt->stream;
    //
    // A second host thread issues the tape and adjoint streams' work from a queue
    // of steps, so that a full forward queue, which blocks the thread issuing
    // into it, never holds back the work the forward stream may be waiting for.
    struct Job { long n; int tape; };
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Job> jobs;
    bool closing = false;
    long issued = 0;                     // jobs whose work the second thread has issued
    wedg_status worker_status = WEDG_OK;
    std::vector<long> tape_job(ntapes, -1);   // the job that last used each tape
    int device = 0;
    HIPOK(hipGetDevice(&device));
    auto issue = [&](const Job& jb) -> wedg_status {
        AdjTape& tp = A->tape[jb.tape];
        HIPOK(hipStreamWaitEvent(ts, tp.copied, 0));
        wedg_status s2;
        if ((s2 = tape_step(A->wt, tp, tt[jb.n], dt, jb.n + 1))) return s2;
        HIPOK(hipEventRecord(tp.taped, ts));
        HIPOK(hipStreamWaitEvent(as, tp.taped, 0));
        if ((s2 = adjoint_step(A, tp, jb.n, dt))) return s2;
        HIPOK(hipEventRecord(tp.used, as));
        return WEDG_OK;
    };
    std::thread worker([&]() {
        if (hipSetDevice(device) != hipSuccess) worker_status = WEDG_ERR_HIP;
        for (;;) {
            Job jb;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return closing || !jobs.empty(); });
                if (jobs.empty()) return;
                jb = jobs.front();
                jobs.pop_front();
            }
            wedg_status s2 = worker_status == WEDG_OK ? issue(jb) : worker_status;
            {
                std::lock_guard<std::mutex> lk(mu);
                if (s2 != WEDG_OK && worker_status == WEDG_OK) worker_status = s2;
                ++issued;
            }
            cv.notify_all();
        }
    });
    long pushed = 0;
    auto leaf = [&](long n) -> wedg_status {
        const int ti = (int)(n % ntapes);
        AdjTape& tp = A->tape[ti];
        if (tape_job[ti] >= 0) {
            // the tape's last use must have been issued (its `used` recorded) before the wait
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return issued > tape_job[ti]; });
            lk.unlock();
            HIPOK(hipStreamWaitEvent(fs, tp.used, 0));
        }
        HIPOK(copy_state(cur_u, tp.Yu[0], nU, cur_v, tp.Yv[0], nV, fs));
        HIPOK(hipEventRecord(tp.copied, fs));
        ++A->stats.taped_steps;
        tape_job[ti] = pushed;
        {
            std::lock_guard<std::mutex> lk(mu);
            jobs.push_back(Job{n, ti});
            ++pushed;
        }
        cv.notify_all();
        return WEDG_OK;
    };
    std::vector<int> free_slots;
    for (int s = nsnap - 1; s >= 1; --s) free_slots.push_back(s);
    int max_used = 1;
    std::function<wedg_status(long, long, int)> reverse = [&](long a, long b, int slot_a) -> wedg_status {
        wedg_status s2;
        if (b - a == 1) return leaf(a);
        const int sfree = (int)free_slots.size();
        if (sfree == 0) {
            for (long i = b - 1; i >= a; --i) {
                restore(slot_a);
                if ((s2 = advance(a, i, -1))) return s2;
                if ((s2 = leaf(i))) return s2;
            }
            return WEDG_OK;
        }
        const long m = a + revolve_split(b - a, sfree);
        const int sm = free_slots.back();
        free_slots.pop_back();
        max_used = std::max(max_used, nsnap - (int)free_slots.size());
        if ((s2 = advance(a, m, sm))) return s2;
        if ((s2 = reverse(m, b, sm))) return s2;
        free_slots.push_back(sm);
        restore(slot_a);
        return reverse(a, m, slot_a);
    };
    const auto c0 = std::chrono::steady_clock::now();
    HIPOK(hipMemsetAsync(snap_u(0), 0, sz * sizeof(double), fs));   // y_0 = 0 in slot 0
    ++A->stats.snapshots_taken;
    st = reverse(0, nsteps, 0);
    {
        std::lock_guard<std::mutex> lk(mu);
        closing = true;
    }
    cv.notify_all();
    worker.join();
    if (st) return st;
    if (worker_status) return worker_status;
    A->stats.snapshots_used = max_used;
    const auto c1 = std::chrono::steady_clock::now();
    HIPOK(hipStreamSynchronize(as));
    HIPOK(hipStreamSynchronize(fs));
    const auto c2 = std::chrono::steady_clock::now();
    A->stats.enqueue_ms = std::chrono::duration<double, std::milli>(c1 - c0).count();
    A->stats.total_ms = std::chrono::duration<double, std::milli>(c2 - c0).count();
    if ((st = collect(ws, 1)) || (st = collect(A->wt, 1))) return st;
    if ((st = read_fail(A))) return st;
    if ((st = sum_misfit(A, nsteps, J))) return st;
    if ((st = d2h(A, dJ_dc2, A->gc2, A->E))) return st;
    return d2h(A, dJ_dbeta, A->gbe, A->E);
}

extern "C" wedg_status wedg_adj_last_stats(const wedg_adj* A, wedg_adj_stats* st) {
    if (!A || !st) return WEDG_ERR_ARGUMENT;
    *st = A->stats;
    return WEDG_OK;
}

// F(Y) of one state into tape set 0 (Y1 = the state, K1 = F(Y)).
static wedg_status rhs_into_tape(wedg_adj* A, const double* u, const double* v, double t) {
    AdjTape& tp = A->tape[0];
    wedg_status st;
    if ((st = h2d(A, tp.Yu[0], u, A->nU())) || (st = h2d(A, tp.Yv[0], v, A->nV()))) return st;
    if ((st = reset_err(A->ws))) return st;
    KArgs a = stage_args(A->ws, 0, tp.Yu[0], tp.Yv[0], tp.Yu[0], tp.Yv[0], tp.Ku[0], tp.Kv[0], t, 0.0, 0, false, t);
    if ((st = run_kernel(A->ws, a, 0)) || (st = run_kernel(A->ws, a, 1))) return st;
    return collect(A->ws, 0);
}

extern "C" wedg_status wedg_adj_rhs_vjp(wedg_adj* A, const double* u, const double* v, double t, const double* ub,
                                        const double* vb, double* u_bar, double* v_bar, double* c2_bar,
                                        double* beta_bar) {
    if (!A || !u || !v || !ub || !vb || !u_bar || !v_bar || !c2_bar || !beta_bar) return WEDG_ERR_ARGUMENT;
    wedg_status st;
    if ((st = rhs_into_tape(A, u, v, t))) return st;
    if ((st = h2d(A, A->kbu, ub, A->nU())) || (st = h2d(A, A->kbv, vb, A->nV()))) return st;
    const hipStream_t s = A->ws->stream;
    HIPOK(hipMemsetAsync(A->fail, 0, sizeof(int), s));
    if ((st = zero_dev(s, A->gc2, A->E)) || (st = zero_dev(s, A->gbe, A->E))) return st;
    AdjArgs a = adj_args(A);
    AdjTape& tp = A->tape[0];
    a.stage = 0;
    a.yu = tp.Yu[0]; a.yv = tp.Yv[0]; a.ku = tp.Ku[0]; a.kv = tp.Kv[0];
    a.kbu = A->kbu; a.kbv = A->kbv;
    HIPOK(launch_adj(a, A->qu, A->nv, 0, s));
    HIPOK(launch_adj(a, A->qu, A->nv, 1, s));
    if ((st = read_fail(A))) return st;
    if ((st = d2h(A, u_bar, A->accu, A->nU())) || (st = d2h(A, v_bar, A->accv, A->nV()))) return st;
    if ((st = d2h(A, c2_bar, A->gc2, A->E))) return st;
    return d2h(A, beta_bar, A->gbe, A->E);
}

extern "C" wedg_status wedg_adj_rhs_tl(wedg_adj* A, const double* u, const double* v, double t, const double* du,
                                       const double* dv, const double* dc2, const double* dbeta, double* dut,
                                       double* dvt) {
    if (!A || !u || !v || !du || !dv || !dc2 || !dbeta || !dut || !dvt) return WEDG_ERR_ARGUMENT;
    wedg_status st;
    if ((st = rhs_into_tape(A, u, v, t))) return st;
    AdjTape& tp = A->tape[0];
    if ((st = h2d(A, tp.Yu[1], du, A->nU())) || (st = h2d(A, tp.Yv[1], dv, A->nV()))) return st;
    if ((st = h2d(A, A->tdc2, dc2, A->E)) || (st = h2d(A, A->tdbe, dbeta, A->E))) return st;
    const hipStream_t s = A->ws->stream;
    HIPOK(hipMemsetAsync(A->fail, 0, sizeof(int), s));
    AdjArgs a = adj_args(A);
    a.yu = tp.Yu[0]; a.yv = tp.Yv[0]; a.ku = tp.Ku[0]; a.kv = tp.Kv[0];
    a.du = tp.Yu[1]; a.dv = tp.Yv[1]; a.dc2 = A->tdc2; a.dbe = A->tdbe;
    a.outu = tp.Yu[2]; a.outv = tp.Yv[2];
    HIPOK(launch_adj(a, A->qu, A->nv, 2, s));
    if ((st = read_fail(A))) return st;
    if ((st = d2h(A, dut, tp.Yu[2], A->nU()))) return st;
    return d2h(A, dvt, tp.Yv[2], A->nV());
}

// ---------------------------------------------------------------------------
// Batched gradients (wedg_adj_batch_*): several instances run their binomial
// checkpointing schedules in lockstep rounds on one stream. In a round every
// instance with work left takes its next unit, a forward step or the reversal
// of a step (the step re-run with its stages taped, then its adjoint); per group
// of instances with the same (q_u, q_v, kernel layout) a round is one k_prep and
// four k_stage launches over every instance of the group, then k_misfit,
// k_adj_a, three k_adj_ba and k_adj_b over the instances reversing a step. The
// arguments of a round (stage and adjoint arguments, block maps) reach the
// device in one copy through a ring of pinned buffers. Every instance computes
// exactly what its own wedg_adj_gradient computes.

struct wedg_adj_batch {
    std::vector<wedg_adj*> inst;
    // Each degree group's rounds run on stream[g % kStreams]: the groups share
    // no buffers, so one group's launches fill the SIMDs another's leave idle.
    // In a gradient the reversed steps' adjoint launches of group g run on
    // astream[g % kStreams], one round behind its forward launches: a reversal
    // tapes into the tape of its round's parity and copies its start state into
    // it, so the next round's forward launches touch nothing it reads.
    static constexpr int kStreams = 2;
    hipStream_t stream[kStreams] = {};
    hipStream_t astream[kStreams] = {};
    static constexpr int kRing = 8;
    char* hbuf[kRing] = {};
    char* dbuf[kRing] = {};
    hipEvent_t done[kRing][2 * kStreams] = {};   // a round's launches on each of the streams
    bool used[kRing] = {};
    hipEvent_t fdone[kStreams] = {};             // a group's forward launches of the current round
    hipEvent_t adone[2][kStreams] = {};          // the adjoint launches of the last round of each parity
    hipEvent_t join = nullptr;
    size_t cap = 0;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    double device_ms = 0.0, enqueue_ms = 0.0;
};

namespace {

hipError_t launch_fwd_batch(const BatchArgs& b, int grid, int qu, int nv, int which, hipStream_t s, int mode,
                            bool two) {
    if (two) {
        switch (qu) {
            case 1: return launch_batch2_q1(b, grid, nv, which, s, mode);
            case 2: return launch_batch2_q2(b, grid, nv, which, s, mode);
            case 3: return launch_batch2_q3(b, grid, nv, which, s, mode);
            case 4: return launch_batch2_q4(b, grid, nv, which, s, mode);
            case 5: return launch_batch2_q5(b, grid, nv, which, s, mode);
            case 6: return launch_batch2_q6(b, grid, nv, which, s, mode);
            default: return hipErrorInvalidValue;
        }
    }
    switch (qu) {
        case 1: return launch_batch_q1(b, grid, nv, which, s, mode);
        case 2: return launch_batch_q2(b, grid, nv, which, s, mode);
        case 3: return launch_batch_q3(b, grid, nv, which, s, mode);
        case 4: return launch_batch_q4(b, grid, nv, which, s, mode);
        case 5: return launch_batch_q5(b, grid, nv, which, s, mode);
        case 6: return launch_batch_q6(b, grid, nv, which, s, mode);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_adj_batch(const BatchArgs& b, int grid, int qu, int nv, int which, hipStream_t s) {
    switch (qu) {
        case 1: return launch_adjb_q1(b, grid, nv, which, s);
        case 2: return launch_adjb_q2(b, grid, nv, which, s);
        case 3: return launch_adjb_q3(b, grid, nv, which, s);
        case 4: return launch_adjb_q4(b, grid, nv, which, s);
        case 5: return launch_adjb_q5(b, grid, nv, which, s);
        case 6: return launch_adjb_q6(b, grid, nv, which, s);
        default: return hipErrorInvalidValue;
    }
}

// One unit of an instance's schedule: the forward step n from state `in` into
// state `out`, or (rev) the reversal of step n from state `in`. States: -1 the
// working state of the instance's workspace, s >= 0 snapshot slot s.
struct AdjUnit {
    long n;
    int rev, in, out;
};

// The binomial checkpointing schedule of wedg_adj_gradient as units, in the
// order it issues them, with its counts.
void revolve_units(long nsteps, int nsnap, std::vector<AdjUnit>& units, wedg_adj_stats& stats) {
    units.clear();
    int cur = 0;
    std::vector<int> free_slots;
    for (int s = nsnap - 1; s >= 1; --s) free_slots.push_back(s);
    int max_used = 1;
    auto advance = [&](long a, long m, int save) {
        for (long i = a; i < m; ++i) {
            const int out = (i == m - 1 && save >= 0) ? save : -1;
            units.push_back({i, 0, cur, out});
            cur = out;
            ++stats.forward_steps;
        }
        if (save >= 0) ++stats.snapshots_taken;
    };
    auto leaf = [&](long n) {
        units.push_back({n, 1, cur, -1});
        ++stats.taped_steps;
        ++stats.adjoint_steps;
    };
    std::function<void(long, long, int)> reverse = [&](long a, long b, int slot_a) {
        if (b - a == 1) {
            leaf(a);
            return;
        }
        const int sfree = (int)free_slots.size();
        if (sfree == 0) {
            for (long i = b - 1; i >= a; --i) {
                cur = slot_a;
                advance(a, i, -1);
                leaf(i);
            }
            return;
        }
        const long m = a + revolve_split(b - a, sfree);
        const int sm = free_slots.back();
        free_slots.pop_back();
        max_used = std::max(max_used, nsnap - (int)free_slots.size());
        advance(a, m, sm);
        reverse(m, b, sm);
        free_slots.push_back(sm);
        cur = slot_a;
        reverse(a, m, slot_a);
    };
    ++stats.snapshots_taken;   // y_0 in slot 0
    reverse(0, nsteps, 0);
    stats.snapshots_used = max_used;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Bytes of one round's arguments when every instance takes part in every launch.
size_t round_bytes(const wedg_adj_batch* b) {
    size_t n = 0, blocks = 0, mblocks = 0;
    for (wedg_adj* A : b->inst) {
        ++n;
        blocks += (A->E + kPairWave - 1) / kPairWave;   // the forward launches' most (two lanes)
        mblocks += (A->N + kWave - 1) / kWave;
    }
    const size_t per_group_arrays = 4 + 6 + 6;
    return 4 * align256(n * sizeof(KArgs)) + 6 * align256(2 * n * sizeof(AdjArgs)) +
           2 * align256(blocks * sizeof(int)) + align256(mblocks * sizeof(int)) + 3 * align256(n * sizeof(int)) +
           align256(n * sizeof(CopyJob)) + b->inst.size() * (per_group_arrays + 1) * 256 + 4096;
}

// the group key of an instance: kernels are instantiated per (q_u, q_v, layout)
long group_key(const wedg_adj* A) { return (((long)A->qu * 16 + A->nv) * 16 + A->p->mode) * 2 + (A->ws->two() ? 1 : 0); }

// Blocks (waves) of an instance's forward stage launch: 64 elements per wave,
// 32 with two lanes per element.
int fwd_blocks(const wedg_adj* A) {
    const int w = A->ws->two() ? kPairWave : kWave;
    return (A->E + w - 1) / w;
}

wedg_status batch_prepare_instance(wedg_adj* A, long nsteps, int nsnap, bool gradient) {
    wedg_workspace* ws = A->ws;
    const hipStream_t s = ws->stream;
    wedg_status st;
    if ((st = jp_capacity(A, nsteps))) return st;
    const size_t nU = A->nU(), nV = A->nV(), sz = nU + nV;
    if (gradient && nsnap > A->nsnap_cap) {
        HIPOK(hipStreamSynchronize(s));
        if (A->snaps) (void)hipFree(A->snaps);
        A->snaps = nullptr;
        A->nsnap_cap = 0;
        HIPOK(hipMalloc((void**)&A->snaps, (size_t)nsnap * sz * sizeof(double)));
        A->nsnap_cap = nsnap;
    }
    std::memset(&A->stats, 0, sizeof(A->stats));
    if (gradient && (st = ensure_tapes(A, 2))) return st;   // a tape per round parity
    if ((st = reset_err(ws))) return st;
    HIPOK(hipMemsetAsync(A->fail, 0, sizeof(int), s));
    HIPOK(hipMemsetAsync(ws->yu, 0, nU * sizeof(double), s));
    HIPOK(hipMemsetAsync(ws->yv, 0, nV * sizeof(double), s));
    if (gradient) {
        if ((st = zero_dev(s, A->gc2, A->E)) || (st = zero_dev(s, A->gbe, A->E)) || (st = zero_dev(s, A->ybu, nU)) ||
            (st = zero_dev(s, A->ybv, nV)))
            return st;
        HIPOK(hipMemsetAsync(A->snaps, 0, sz * sizeof(double), s));   // y_0 = 0 in slot 0
    }
    HIPOK(hipStreamSynchronize(s));
    return WEDG_OK;
}

// Runs the units of every instance in lockstep rounds (units[i] of instance i).
// forward_misfit: every unit is a forward step and forms the misfit of its end
// state; otherwise the misfit is formed by the reversals (as the gradient does).
wedg_status batch_rounds(wedg_adj_batch* B, const std::vector<std::vector<AdjUnit>>& units, const double* dt,
                         bool forward_misfit) {
    const int n = (int)B->inst.size();
    constexpr int NS = wedg_adj_batch::kStreams;
    // groups
    std::vector<long> keys;
    for (wedg_adj* A : B->inst) keys.push_back(group_key(A));
    std::vector<long> gk = keys;
    std::sort(gk.begin(), gk.end());
    gk.erase(std::unique(gk.begin(), gk.end()), gk.end());
    std::vector<std::vector<int>> gmem(gk.size());
    for (int i = 0; i < n; ++i) gmem[std::lower_bound(gk.begin(), gk.end(), keys[i]) - gk.begin()].push_back(i);
    // times after n steps, accumulated as the forward loop does (src/solver.jl:213-216)
    std::vector<std::vector<double>> tt(n);
    long rounds = 0;
    for (int i = 0; i < n; ++i) {
        long ns = 0;
        for (const AdjUnit& u : units[i]) ns = std::max(ns, u.n + 1);
        tt[i].assign(ns + 1, 0.0);
        for (long k = 0; k < ns; ++k) tt[i][k + 1] = tt[i][k] + dt[i];
        rounds = std::max(rounds, (long)units[i].size());
    }
    auto state_u = [&](wedg_adj* A, int id) -> double* {
        return id < 0 ? A->ws->yu : A->snaps + (size_t)id * (A->nU() + A->nV());
    };
    auto state_v = [&](wedg_adj* A, int id) -> double* {
        return id < 0 ? A->ws->yv : A->snaps + (size_t)id * (A->nU() + A->nV()) + A->nU();
    };
    HIPOK(hipEventRecord(B->t0, B->stream[0]));
    for (int k = 0; k < NS; ++k) {
        if (k > 0) HIPOK(hipStreamWaitEvent(B->stream[k], B->t0, 0));
        HIPOK(hipStreamWaitEvent(B->astream[k], B->t0, 0));
        HIPOK(hipEventRecord(B->adone[0][k], B->astream[k]));
        HIPOK(hipEventRecord(B->adone[1][k], B->astream[k]));
    }
    const auto c0 = std::chrono::steady_clock::now();
    for (long r = 0; r < rounds; ++r) {
        const int ring = (int)(r % wedg_adj_batch::kRing);
        const int par = (int)(r & 1);   // the tape of this round's reversals
        if (B->used[ring])
            for (int k = 0; k < 2 * NS; ++k) HIPOK(hipEventSynchronize(B->done[ring][k]));
        char* hb = B->hbuf[ring];
        size_t off = 0;
        auto take = [&](size_t bytes) {
            const size_t o = off;
            off = align256(off + bytes);
            return o;
        };
        struct Launch {
            int qu, nv, mode, which, grid, adj;   // adj: 1 adjoint kernel, 2 the start-state copies
            size_t args, binst, bfirst;
            bool two;    // forward launches: the two-lane kernels
            bool back;   // on the group's adjoint stream (the gradient's misfit and adjoint kernels)
        };
        std::vector<Launch> launches;
        struct Range {
            size_t begin, end, first, count;   // the group's argument bytes and its launches
            int stream;
        };
        std::vector<Range> ranges;
        for (size_t g = 0; g < gk.size(); ++g) {
            std::vector<int> F, R;
            for (int i : gmem[g])
                if (r < (long)units[i].size()) {
                    F.push_back(i);
                    if (units[i][r].rev) R.push_back(i);
                }
            if (F.empty()) continue;
            ranges.push_back({off, off, launches.size(), 0, (int)(g % NS)});
            wedg_adj* A0 = B->inst[F[0]];
            const int qu = A0->qu, nv = A0->nv, mode = A0->p->mode;
            // forward: block map and the four stages' arguments
            const bool two = A0->ws->two();
            int fblocks = 0;
            for (int i : F) fblocks += fwd_blocks(B->inst[i]);
            const size_t o_bi = take(fblocks * sizeof(int)), o_bf = take(F.size() * sizeof(int));
            {
                int* bi = (int*)(hb + o_bi);
                int* bf = (int*)(hb + o_bf);
                int blk = 0;
                for (size_t x = 0; x < F.size(); ++x) {
                    const int nb = fwd_blocks(B->inst[F[x]]);
                    bf[x] = blk;
                    for (int k = 0; k < nb; ++k) bi[blk++] = (int)x;
                }
            }
            size_t o_st[4];
            for (int sg = 0; sg < 4; ++sg) o_st[sg] = take(F.size() * sizeof(KArgs));
            for (size_t x = 0; x < F.size(); ++x) {
                wedg_adj* A = B->inst[F[x]];
                wedg_workspace* ws = A->ws;
                const AdjUnit& u = units[F[x]][r];
                const double t = tt[F[x]][u.n], d = dt[F[x]], hd = 0.5 * d, th = t + hd, tf = t + d;
                const double* iu = state_u(A, u.in);
                const double* iv = state_v(A, u.in);
                const long step = u.n + 1;
                KArgs a[4];
                if (!u.rev) {
                    double* ou = state_u(A, u.out);
                    double* ov = state_v(A, u.out);
                    a[0] = stage_args(ws, 1, iu, iv, iu, iv, ws->su, ws->sv, t, hd, step, true, th);
                    a[1] = stage_args(ws, 2, ws->su, ws->sv, iu, iv, ws->ou, ws->ov, th, hd, step, true, th);
                    a[2] = stage_args(ws, 3, ws->ou, ws->ov, iu, iv, ws->su, ws->sv, th, d, step, true, tf);
                    a[3] = stage_args(ws, 4, ws->su, ws->sv, iu, iv, ou, ov, tf, d / 6.0, step, false, tf);
                } else {
                    AdjTape& tp = A->tape[par];
                    const double ts[4] = {t, th, th, tf}, cs[4] = {hd, hd, d, d / 6.0}, tn[4] = {th, th, tf, tf};
                    for (int sg = 0; sg < 4; ++sg) {
                        const double* yu = sg == 0 ? iu : tp.Yu[sg];
                        const double* yv = sg == 0 ? iv : tp.Yv[sg];
                        double* ou = sg < 3 ? tp.Yu[sg + 1] : tp.ynu;
                        double* ov = sg < 3 ? tp.Yv[sg + 1] : tp.ynv;
                        a[sg] = stage_args(ws, sg + 1, yu, yv, iu, iv, ou, ov, ts[sg], cs[sg], step, sg < 3, tn[sg]);
                        a[sg].recu = tp.Ku[sg];
                        a[sg].recv = tp.Kv[sg];
                    }
                }
                for (int sg = 0; sg < 4; ++sg) std::memcpy(hb + o_st[sg] + x * sizeof(KArgs), &a[sg], sizeof(KArgs));
                if (u.rev) ++A->stats.taped_steps;
                else ++A->stats.forward_steps;
            }
            if (!forward_misfit && !R.empty()) {
                // the reversed steps' start states into their tapes, for the adjoint kernels
                const size_t o_cj = take(R.size() * sizeof(CopyJob));
                CopyJob* cj = (CopyJob*)(hb + o_cj);
                for (size_t x = 0; x < R.size(); ++x) {
                    wedg_adj* A = B->inst[R[x]];
                    const AdjUnit& u = units[R[x]][r];
                    AdjTape& tp = A->tape[par];
                    cj[x] = CopyJob{state_u(A, u.in), tp.Yu[0], state_v(A, u.in), tp.Yv[0], A->nU(), A->nV()};
                }
                launches.push_back({qu, nv, mode, 0, (int)R.size(), 2, o_cj, 0, 0, false, false});
            }
            launches.push_back({qu, nv, mode, 0, fblocks, 0, o_st[0], o_bi, o_bf, two, false});   // k_prep
            for (int sg = 0; sg < 4; ++sg)
                launches.push_back({qu, nv, mode, 1, fblocks, 0, o_st[sg], o_bi, o_bf, two, false});
            // misfit of the forward steps' end states (forward runs)
            std::vector<int> M = forward_misfit ? F : R;
            if (!M.empty()) {
                int mblocks = 0;
                for (int i : M) mblocks += (B->inst[i]->N + kWave - 1) / kWave;
                const size_t o_mi = take(mblocks * sizeof(int)), o_mf = take(M.size() * sizeof(int));
                const size_t o_ma = take(2 * M.size() * sizeof(AdjArgs));
                int* bi = (int*)(hb + o_mi);
                int* bf = (int*)(hb + o_mf);
                int blk = 0;
                for (size_t x = 0; x < M.size(); ++x) {
                    const int nb = (B->inst[M[x]]->N + kWave - 1) / kWave;
                    bf[x] = blk;
                    for (int k = 0; k < nb; ++k) bi[blk++] = (int)x;
                }
                for (size_t x = 0; x < M.size(); ++x) {
                    wedg_adj* A = B->inst[M[x]];
                    const AdjUnit& u = units[M[x]][r];
                    AdjArgs mis = adj_args(A);
                    const size_t row = (size_t)A->N * A->nu;
                    if (forward_misfit) {
                        mis.yu = state_u(A, u.out);
                        mis.tgt = A->dtgt + (size_t)u.n * row;
                        mis.jp = A->djp + (size_t)u.n * row;
                        mis.ybu = nullptr;
                    } else {
                        AdjArgs st5[5];
                        const AdjTape& tp = A->tape[par];
                        adjoint_step_args(A, tp, u.n, dt[M[x]], tp.Yu[0], tp.Yv[0], mis, st5);
                    }
                    std::memcpy(hb + o_ma + (2 * x) * sizeof(AdjArgs), &mis, sizeof(AdjArgs));
                    std::memcpy(hb + o_ma + (2 * x + 1) * sizeof(AdjArgs), &mis, sizeof(AdjArgs));
                }
                launches.push_back({qu, nv, mode, 3, mblocks, 1, o_ma, o_mi, o_mf, false, !forward_misfit});
            }
            ranges.back().end = off;
            ranges.back().count = launches.size() - ranges.back().first;
            if (R.empty() || forward_misfit) continue;
            // the adjoint of the reversed steps
            int ablocks = 0;
            for (int i : R) ablocks += (B->inst[i]->E + kWave - 1) / kWave;
            const size_t o_ai = take(ablocks * sizeof(int)), o_af = take(R.size() * sizeof(int));
            {
                int* bi = (int*)(hb + o_ai);
                int* bf = (int*)(hb + o_af);
                int blk = 0;
                for (size_t x = 0; x < R.size(); ++x) {
                    const int nb = (B->inst[R[x]]->E + kWave - 1) / kWave;
                    bf[x] = blk;
                    for (int k = 0; k < nb; ++k) bi[blk++] = (int)x;
                }
            }
            size_t o_ad[5];   // k_adj_a(4), k_adj_ba(4, 3), (3, 2), (2, 1), k_adj_b(1)
            for (int k = 0; k < 5; ++k) o_ad[k] = take(2 * R.size() * sizeof(AdjArgs));
            for (size_t x = 0; x < R.size(); ++x) {
                wedg_adj* A = B->inst[R[x]];
                const AdjUnit& u = units[R[x]][r];
                AdjArgs mis, st5[5];
                const AdjTape& tp = A->tape[par];
                adjoint_step_args(A, tp, u.n, dt[R[x]], tp.Yu[0], tp.Yv[0], mis, st5);
                const AdjArgs pairs[5][2] = {{st5[4], st5[4]}, {st5[4], st5[3]}, {st5[3], st5[2]}, {st5[2], st5[1]},
                                             {st5[1], st5[1]}};
                for (int k = 0; k < 5; ++k)
                    std::memcpy(hb + o_ad[k] + 2 * x * sizeof(AdjArgs), pairs[k], 2 * sizeof(AdjArgs));
                ++A->stats.adjoint_steps;
            }
            launches.push_back({qu, nv, mode, 0, ablocks, 1, o_ad[0], o_ai, o_af, false, true});
            for (int k = 1; k <= 3; ++k)
                launches.push_back({qu, nv, mode, 2, ablocks, 1, o_ad[k], o_ai, o_af, false, true});
            launches.push_back({qu, nv, mode, 1, ablocks, 1, o_ad[4], o_ai, o_af, false, true});
            ranges.back().end = off;
            ranges.back().count = launches.size() - ranges.back().first;
        }
        if (off > B->cap) return WEDG_ERR_ARGUMENT;
        char* db = B->dbuf[ring];
        // a reversal tapes into the tape of this round's parity: the forward
        // stream waits for the adjoint launches of two rounds ago, which read it
        if (!forward_misfit)
            for (int k = 0; k < NS; ++k) HIPOK(hipStreamWaitEvent(B->stream[k], B->adone[par][k], 0));
        for (const Range& G : ranges) {
            const hipStream_t gs = B->stream[G.stream], as = B->astream[G.stream];
            HIPOK(hipMemcpyAsync(db + G.begin, hb + G.begin, G.end - G.begin, hipMemcpyHostToDevice, gs));
            bool joined = false;
            for (size_t x = G.first; x < G.first + G.count; ++x) {
                const Launch& L = launches[x];
                if (L.back && !joined) {   // the group's forward launches of this round come first
                    HIPOK(hipEventRecord(B->fdone[G.stream], gs));
                    HIPOK(hipStreamWaitEvent(as, B->fdone[G.stream], 0));
                    joined = true;
                }
                const hipStream_t ls = L.back ? as : gs;
                if (L.adj == 2) {
                    k_copy_states<<<dim3(256, L.grid), 256, 0, ls>>>((const CopyJob*)(const void*)(db + L.args));
                    HIPOK(hipGetLastError());
                    continue;
                }
                BatchArgs ba;
                ba.args = (const KArgs*)(const void*)(db + L.args);
                ba.binst = (const int*)(db + L.binst);
                ba.bfirst = (const int*)(db + L.bfirst);
                if (L.adj) HIPOK(launch_adj_batch(ba, L.grid, L.qu, L.nv, L.which, ls));
                else HIPOK(launch_fwd_batch(ba, L.grid, L.qu, L.nv, L.which, ls, L.mode, L.two));
            }
        }
        for (int k = 0; k < NS; ++k) {
            HIPOK(hipEventRecord(B->done[ring][k], B->stream[k]));
            HIPOK(hipEventRecord(B->done[ring][NS + k], B->astream[k]));
            if (!forward_misfit) HIPOK(hipEventRecord(B->adone[par][k], B->astream[k]));
        }
        B->used[ring] = true;
    }
    const auto c1 = std::chrono::steady_clock::now();
    for (int k = 0; k < NS; ++k) {
        if (k > 0) {
            HIPOK(hipEventRecord(B->join, B->stream[k]));
            HIPOK(hipStreamWaitEvent(B->stream[0], B->join, 0));
        }
        HIPOK(hipEventRecord(B->join, B->astream[k]));
        HIPOK(hipStreamWaitEvent(B->stream[0], B->join, 0));
    }
    HIPOK(hipEventRecord(B->t1, B->stream[0]));
    HIPOK(hipStreamSynchronize(B->stream[0]));
    float ms = 0.f;
    B->device_ms = hipEventElapsedTime(&ms, B->t0, B->t1) == hipSuccess ? ms : 0.0;
    B->enqueue_ms = std::chrono::duration<double, std::milli>(c1 - c0).count();
    return WEDG_OK;
}

}  // namespace

extern "C" wedg_status wedg_adj_batch_create(wedg_adj* const* a, int n, wedg_adj_batch** out) {
    if (!a || n < 1 || !out) return WEDG_ERR_ARGUMENT;
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (!a[i]) return WEDG_ERR_ARGUMENT;
        for (int k = 0; k < i; ++k)
            if (a[k] == a[i]) return WEDG_ERR_ARGUMENT;
    }
    auto* B = new wedg_adj_batch();
    B->inst.assign(a, a + n);
    auto fail = [&](wedg_status s) {
        wedg_adj_batch_destroy(B);
        return s;
    };
    for (int k = 0; k < wedg_adj_batch::kStreams; ++k) {
        if (hipStreamCreateWithFlags(&B->stream[k], hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&B->astream[k], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&B->fdone[k], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&B->adone[0][k], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&B->adone[1][k], hipEventDisableTiming) != hipSuccess)
            return fail(WEDG_ERR_HIP);
    }
    if (hipEventCreate(&B->t0) != hipSuccess || hipEventCreate(&B->t1) != hipSuccess) return fail(WEDG_ERR_HIP);
    if (hipEventCreateWithFlags(&B->join, hipEventDisableTiming) != hipSuccess) return fail(WEDG_ERR_HIP);
    B->cap = round_bytes(B);
    for (int k = 0; k < wedg_adj_batch::kRing; ++k) {
        if (hipHostMalloc((void**)&B->hbuf[k], B->cap, hipHostMallocDefault) != hipSuccess) return fail(WEDG_ERR_HIP);
        if (hipMalloc((void**)&B->dbuf[k], B->cap) != hipSuccess) return fail(WEDG_ERR_HIP);
        for (auto& ev : B->done[k])
            if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return fail(WEDG_ERR_HIP);
    }
    *out = B;
    return WEDG_OK;
}

extern "C" wedg_status wedg_adj_batch_destroy(wedg_adj_batch* B) {
    if (!B) return WEDG_OK;
    for (auto& st : B->stream)
        if (st) (void)hipStreamSynchronize(st);
    for (auto& st : B->astream)
        if (st) (void)hipStreamSynchronize(st);
    for (int k = 0; k < wedg_adj_batch::kRing; ++k) {
        if (B->hbuf[k]) (void)hipHostFree(B->hbuf[k]);
        if (B->dbuf[k]) (void)hipFree(B->dbuf[k]);
        for (auto& ev : B->done[k])
            if (ev) (void)hipEventDestroy(ev);
    }
    if (B->t0) (void)hipEventDestroy(B->t0);
    if (B->t1) (void)hipEventDestroy(B->t1);
    if (B->join) (void)hipEventDestroy(B->join);
    for (int k = 0; k < wedg_adj_batch::kStreams; ++k) {
        if (B->fdone[k]) (void)hipEventDestroy(B->fdone[k]);
        if (B->adone[0][k]) (void)hipEventDestroy(B->adone[0][k]);
        if (B->adone[1][k]) (void)hipEventDestroy(B->adone[1][k]);
    }
    for (auto& st : B->stream)
        if (st) (void)hipStreamDestroy(st);
    for (auto& st : B->astream)
        if (st) (void)hipStreamDestroy(st);
    delete B;
    return WEDG_OK;
}

// status[i] (or the return value) reports instance i's failure; the others finish.
static wedg_status batch_finish(wedg_adj_batch* B, const long* nsteps, double* J, int* status, bool gradient,
                                double* const* dJ_dc2, double* const* dJ_dbeta) {
    wedg_status first = WEDG_OK;
    for (size_t i = 0; i < B->inst.size(); ++i) {
        wedg_adj* A = B->inst[i];
        wedg_status si = collect(A->ws, 1);
        if (si == WEDG_OK && gradient) si = read_fail(A);
        if (si == WEDG_OK) si = sum_misfit(A, nsteps[i], &J[i]);
        if (si == WEDG_OK && gradient && dJ_dc2 && dJ_dc2[i]) si = d2h(A, dJ_dc2[i], A->gc2, A->E);
        if (si == WEDG_OK && gradient && dJ_dbeta && dJ_dbeta[i]) si = d2h(A, dJ_dbeta[i], A->gbe, A->E);
        A->stats.enqueue_ms = B->enqueue_ms;
        A->stats.total_ms = B->device_ms;
        if (status) status[i] = si;
        if (si != WEDG_OK && first == WEDG_OK) first = si;
    }
    return first;
}

extern "C" wedg_status wedg_adj_batch_forward(wedg_adj_batch* B, const double* dt, const long* nsteps, double* J,
                                              int* status) {
    if (!B || !dt || !nsteps || !J) return WEDG_ERR_ARGUMENT;
    const size_t n = B->inst.size();
    for (size_t i = 0; i < n; ++i)
        if (nsteps[i] < 1 || !(dt[i] > 0) || B->inst[i]->tgt_steps < nsteps[i]) return WEDG_ERR_ARGUMENT;
    std::vector<std::vector<AdjUnit>> units(n);
    wedg_status st;
    for (size_t i = 0; i < n; ++i) {
        if ((st = batch_prepare_instance(B->inst[i], nsteps[i], 0, false))) return st;
        for (long k = 0; k < nsteps[i]; ++k) units[i].push_back({k, 0, -1, -1});
    }
    if ((st = batch_rounds(B, units, dt, true))) return st;
    return batch_finish(B, nsteps, J, status, false, nullptr, nullptr);
}

extern "C" wedg_status wedg_adj_batch_gradient(wedg_adj_batch* B, const double* dt, const long* nsteps, int nsnap,
                                               double* J, double* const* dJ_dc2, double* const* dJ_dbeta,
                                               int* status) {
    if (!B || !dt || !nsteps || !J || !dJ_dc2 || !dJ_dbeta || nsnap < 1) return WEDG_ERR_ARGUMENT;
    const size_t n = B->inst.size();
    for (size_t i = 0; i < n; ++i)
        if (nsteps[i] < 1 || !(dt[i] > 0) || B->inst[i]->tgt_steps < nsteps[i] || !dJ_dc2[i] || !dJ_dbeta[i])
            return WEDG_ERR_ARGUMENT;
    std::vector<std::vector<AdjUnit>> units(n);
    wedg_status st;
    for (size_t i = 0; i < n; ++i) {
        wedg_adj* A = B->inst[i];
        if ((st = batch_prepare_instance(A, nsteps[i], nsnap, true))) return st;
        wedg_adj_stats plan{};
        revolve_units(nsteps[i], nsnap, units[i], plan);
        A->stats.snapshots_taken = plan.snapshots_taken;
        A->stats.snapshots_used = plan.snapshots_used;
    }
    if ((st = batch_rounds(B, units, dt, false))) return st;
    return batch_finish(B, nsteps, J, status, true, dJ_dc2, dJ_dbeta);
}

extern "C" wedg_status wedg_adj_batch_time(const wedg_adj_batch* B, double* device_ms, double* enqueue_ms) {
    if (!B || !device_ms || !enqueue_ms) return WEDG_ERR_ARGUMENT;
    *device_ms = B->device_ms;
    *enqueue_ms = B->enqueue_ms;
    return WEDG_OK;
}
