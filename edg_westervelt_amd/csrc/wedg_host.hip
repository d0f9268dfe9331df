// Host side of the wedg C ABI (include/wedg.h): problem upload with
// de-duplicated coefficient tables, device-resident workspaces, and the RHS /
// RK4 / advance entry points that launch the per-degree kernels.
#include "wedg_internal.h"

using namespace wedg;

namespace wedg {


uint64_t fnv1a(const double* v, size_t n) {
    uint64_t h = 1469598103934665603ull;
    const unsigned char* b = reinterpret_cast<const unsigned char*>(v);
    for (size_t i = 0; i < n * sizeof(double); ++i) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}


// Eigen-decomposition of a small symmetric matrix A (n x n, column-major) by
// cyclic Jacobi rotations: A = V diag(lam) V^T, eigenvalues ascending.
void sym_eigen(std::vector<double> A, int n, std::vector<double>& lam, std::vector<double>& V) {
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[i + (size_t)n * i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                tot += A[i + (size_t)n * j] * A[i + (size_t)n * j];
                if (i != j) off += A[i + (size_t)n * j] * A[i + (size_t)n * j];
            }
        if (off <= 1e-34 * tot) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[p + (size_t)n * q];
                if (apq == 0.0) continue;
                const double app = A[p + (size_t)n * p], aqq = A[q + (size_t)n * q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < n; ++k) {   // A <- J^T A J
                    const double akp = A[k + (size_t)n * p], akq = A[k + (size_t)n * q];
                    A[k + (size_t)n * p] = c * akp - sn * akq;
                    A[k + (size_t)n * q] = sn * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = A[p + (size_t)n * k], aqk = A[q + (size_t)n * k];
                    A[p + (size_t)n * k] = c * apk - sn * aqk;
                    A[q + (size_t)n * k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[k + (size_t)n * p], vkq = V[k + (size_t)n * q];
                    V[k + (size_t)n * p] = c * vkp - sn * vkq;
                    V[k + (size_t)n * q] = sn * vkp + c * vkq;
                }
            }
    }
    lam.resize(n);
    for (int i = 0; i < n; ++i) lam[i] = A[i + (size_t)n * i];
    // ascending order
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (lam[j] < lam[i]) {
                std::swap(lam[i], lam[j]);
                for (int k = 0; k < n; ++k) std::swap(V[k + (size_t)n * i], V[k + (size_t)n * j]);
            }
}

// Spectral factors of the u-stiffness for constant c^2 on an element
// (src/solver.jl:509-536): with K = S^T diag(w) S the 1D stiffness (its modes
// split by parity, mode 0 its null vector) K = Q diag(lam) Q^T, and the pinned
// stiffness is A = c^2 (I (x) K + K (x) I) + e0 e0^T. Q[a + nu i] keeps each
// eigenvector in a slot of the parity of its modes.
void stiffness_spectrum(const double* S, const double* w, int nu, std::vector<double>& Q,
                        std::vector<double>& lam) {
    const int nq = nu;
    std::vector<double> K((size_t)nu * nu, 0.0);
    for (int a = 0; a < nu; ++a)
        for (int b = 0; b < nu; ++b) {
            double s = 0.0;
            for (int p = 0; p < nq; ++p) s += w[p] * S[p + nq * a] * S[p + nq * b];
            K[a + (size_t)nu * b] = s;
        }
    Q.assign((size_t)nu * nu, 0.0);
    lam.assign(nu, 0.0);
    Q[0] = 1.0;   // the constant mode: S e0 = 0
    for (int par = 0; par < 2; ++par) {
        std::vector<int> idx;
        for (int a = (par == 0 ? 2 : 1); a < nu; a += 2) idx.push_back(a);
        const int m = (int)idx.size();
        if (m == 0) continue;
        std::vector<double> B((size_t)m * m), l, V;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) B[i + (size_t)m * j] = 0.5 * (K[idx[i] + (size_t)nu * idx[j]] + K[idx[j] + (size_t)nu * idx[i]]);
        sym_eigen(B, m, l, V);
        for (int c = 0; c < m; ++c) {
            const int slot = idx[c];
            lam[slot] = l[c];
            for (int i = 0; i < m; ++i) Q[idx[i] + (size_t)nu * slot] = V[i + (size_t)m * c];
        }
    }
}

bool all_equal(const double* p, size_t n) {
    for (size_t i = 1; i < n; ++i)
        if (std::memcmp(&p[i], &p[0], sizeof(double)) != 0) return false;
    return true;
}

}  // namespace wedg

hipError_t wedg::launch(const KArgs& a, wedg_problem* p, int which, hipStream_t s, bool two) {
    if (two) {
        switch (p->qu) {
            case 1: return launch2_q1(a, p->nv, which, s, p->mode);
            case 2: return launch2_q2(a, p->nv, which, s, p->mode);
            case 3: return launch2_q3(a, p->nv, which, s, p->mode);
            case 4: return launch2_q4(a, p->nv, which, s, p->mode);
            case 5: return launch2_q5(a, p->nv, which, s, p->mode);
            case 6: return launch2_q6(a, p->nv, which, s, p->mode);
            default: return hipErrorInvalidValue;
        }
    }
    switch (p->qu) {
        case 1: return launch_q1(a, p->nv, which, s, p->mode);
        case 2: return launch_q2(a, p->nv, which, s, p->mode);
        case 3: return launch_q3(a, p->nv, which, s, p->mode);
        case 4: return launch_q4(a, p->nv, which, s, p->mode);
        case 5: return launch_q5(a, p->nv, which, s, p->mode);
        case 6: return launch_q6(a, p->nv, which, s, p->mode);
        default: return hipErrorInvalidValue;
    }
}

// The split tables of the two-lane kernels (wedg_stage2.h) into p->tab2: per
// lane parity h the values T[p][2k + h] of the 1D tables (0 at the dummy row
// 2k + h >= n of a field with n odd modes, and at the other parity of a middle
// point) and, for one class, its u-solve factors D[2k + h][jj]. The kernels
// read the quadrature weights, W_h and a one-class problem's coefficient
// records from row p in both lanes of a pair, so the problem is admitted only
// when those are mirror-symmetric (row NR-1-p equals row p to 1e-13
// relative); otherwise p->tab2 stays null and the problem runs one lane.
hipError_t wedg::build_split_tables(wedg_problem* p, const SplitSrc& src) {
    const int nu = p->nu, nv = p->nv, nh = p->nh, nq = nu;
    if (nu > 7 || !(p->mode == MODE_UNIFORM || p->mode == MODE_SEPARABLE)) return hipSuccess;
    auto mirror_ok = [](const double* a, int nr, int nc, long stride_c) {
        for (int r = 0; r < nr; ++r)
            for (int c = 0; c < nc; ++c) {
                const double x = a[r + stride_c * c], y = a[nr - 1 - r + stride_c * c];
                if (!(std::fabs(x - y) <= 1e-13 * std::max(std::fabs(x), std::fabs(y)))) return false;
            }
        return true;
    };
    bool sym = mirror_ok(src.w, nq, 1, 0) && mirror_ok(src.W2d_h, nh, nh, nh);
    if (p->mode == MODE_UNIFORM)
        sym = sym && src.cw && src.aw && src.beta && src.d0 && mirror_ok(src.cw, nq, nq, nq) &&
              mirror_ok(src.aw, nq, nq, nq) && mirror_ok(src.beta, nh, nh, nh);
    if (!sym) return hipSuccess;
    const SplitLayout L = split_layout(nu, nv, nh);
    std::vector<double> t2(2 * (size_t)L.n, 0.0);
    // T[pp][2k + hp] of a column-major (nr x ncol) table for a field of n
    // modes; the middle point (nr odd) keeps only the parity `keep`
    auto tv = [&](const double* T, int nr, int pp, int k, int hp, int n, int keep) {
        const int a = 2 * k + hp;
        if (a >= n) return 0.0;
        if ((nr & 1) && pp == nr / 2 && (a & 1) != keep) return 0.0;
        return T[pp + (size_t)nr * a];
    };
    for (int hp = 0; hp < 2; ++hp) {
        double* s2 = t2.data() + (size_t)hp * L.n;
        for (int pp = 0; pp < L.NQS; ++pp) {
            for (int k = 0; k < L.RU; ++k) {
                s2[L.sPu + pp * L.RU + k] = tv(src.P_u, nq, pp, k, hp, nu, 0);   // values: even modes
                s2[L.sSu + pp * L.RU + k] = tv(src.S_u, nq, pp, k, hp, nu, 1);   // derivatives: odd
            }
            if (L.sPv != L.sPu)
                for (int k = 0; k < L.RV; ++k) {
                    s2[L.sPv + pp * L.RV + k] = tv(src.P_u, nq, pp, k, hp, nv, 0);
                    s2[L.sSv + pp * L.RV + k] = tv(src.S_u, nq, pp, k, hp, nv, 1);
                }
        }
        for (int pp = 0; pp < L.NHH; ++pp)
            for (int k = 0; k < L.RV; ++k) s2[L.sPh + pp * L.RV + k] = tv(src.P_h_v, nh, pp, k, hp, nv, 0);
        for (int k = 0; k < L.RU; ++k) {
            const int a = 2 * k + hp;
            const bool ok = a < nu;
            s2[L.sE1lu + k] = ok ? src.E1bl[a] : 0.0;
            s2[L.sE1ru + k] = ok ? src.E1br[a] : 0.0;
            for (int k2 = 0; k2 < L.RU; ++k2) {
                const int a2 = 2 * k2 + hp;
                s2[L.sQ + k2 + L.RU * k] = (ok && a2 < nu) ? p->Q[a2 + (size_t)nu * a] : 0.0;
            }
            if (p->mode == MODE_UNIFORM)
                for (int jj = 0; jj < nu; ++jj) s2[L.sD + k + L.RU * jj] = ok ? src.d0[a + (size_t)nu * jj] : 0.0;
        }
        for (int k = 0; k < L.RV; ++k) {
            const int a = 2 * k + hp;
            const bool ok = a < nv;
            s2[L.sElv + k] = ok ? src.Ebl[a] : 0.0;
            s2[L.sErv + k] = ok ? src.Ebr[a] : 0.0;
            if (L.sE1lv != L.sE1lu) {
                s2[L.sE1lv + k] = ok ? src.E1bl[a] : 0.0;
                s2[L.sE1rv + k] = ok ? src.E1br[a] : 0.0;
            }
        }
    }
    hipError_t he = hipSuccess;
    double* d = p->upload(t2.data(), t2.size(), &he);
    if (he == hipSuccess) p->tab2 = d;
    return he;
}

extern "C" wedg_status wedg_problem_create(const wedg_pre* pre, wedg_problem** out) {
    if (!pre || !out) return WEDG_ERR_ARGUMENT;
    *out = nullptr;
    const int N = pre->N, nu = pre->qu1, nv = pre->qv1, nh = pre->q_high;
    if (N < 1 || nu < 2 || nu > 7 || (nv != nu && nv != nu - 1)) return WEDG_ERR_ARGUMENT;
    if (pre->mass_solver != 0 && pre->mass_solver != 1) return WEDG_ERR_ARGUMENT;
    const int qv = nv - 1;
    int nh_expect = (3 * qv + 1) / 2 + 1;
    if (nh_expect < nu) nh_expect = nu;
    if (nh != nh_expect) return WEDG_ERR_DIMENSION;
    auto* p = new wedg_problem();
    p->stage_bytes = (size_t)4 << 20;
    if (hipHostMalloc(&p->stage, p->stage_bytes, hipHostMallocDefault) != hipSuccess) {
        delete p;
        return WEDG_ERR_HIP;
    }
    p->N = N; p->E = N * N; p->nu = nu; p->nv = nv; p->nh = nh; p->qu = nu - 1;
    p->omega = pre->omega; p->theta = pre->theta; p->tau = pre->tau;
    p->lu = pre->lambda_u; p->lv = pre->lambda_v; p->mass_tol = pre->mass_tol;
    p->mass_cg = pre->mass_solver == 1;
    p->mms_on = pre->mms_on; p->src_on = pre->src_on; p->one_sided = pre->one_sided;
    p->src_omega = pre->src_omega; p->src_Ts = pre->src_Ts;
    for (int s = 0; s < 4; ++s) p->bc[s] = pre->bc[s];
    const long E = p->E;
    const int nq = nu;
    hipError_t he = hipSuccess;
#define UP(dst, src, n)                                  \
    do {                                                 \
        (dst) = p->upload(src, n, &he);                  \
        if (he != hipSuccess) { wedg_problem_destroy(p); return WEDG_ERR_HIP; } \
    } while (0)

    // 1D tables: P_u, S_u, w, Ebl, Ebr, E1bl, E1br, P_h_v, W2d_h (compile-time offsets)
    {
        std::vector<double> t;
        t.insert(t.end(), pre->P_u, pre->P_u + nq * nu);
        t.insert(t.end(), pre->S_u, pre->S_u + nq * nu);
        t.insert(t.end(), pre->w, pre->w + nq);
        t.insert(t.end(), pre->Ebl_u, pre->Ebl_u + nu);
        t.insert(t.end(), pre->Ebr_u, pre->Ebr_u + nu);
        t.insert(t.end(), pre->E1bl_u, pre->E1bl_u + nu);
        t.insert(t.end(), pre->E1br_u, pre->E1br_u + nu);
        t.insert(t.end(), pre->P_h_v, pre->P_h_v + nh * nv);
        t.insert(t.end(), pre->W2d_h, pre->W2d_h + nh * nh);
        std::vector<double> Q, lam;
        stiffness_spectrum(pre->S_u, pre->w, nu, Q, lam);
        t.insert(t.end(), Q.begin(), Q.end());
        // lift rows scaled as in the reference's products (2/h) E[a], (4/h^2) E1[a] (src/rhs.jl:449-463)
        const double h = 1.0 / N, two_h = 2.0 / h, four_hh = 4.0 / (h * h);
        for (int a = 0; a < nu; ++a) t.push_back(two_h * pre->Ebl_u[a]);
        for (int a = 0; a < nu; ++a) t.push_back(two_h * pre->Ebr_u[a]);
        for (int a = 0; a < nu; ++a) t.push_back(four_hh * pre->E1bl_u[a]);
        for (int a = 0; a < nu; ++a) t.push_back(four_hh * pre->E1br_u[a]);
        UP(p->tab, t.data(), t.size());
        p->lam = lam;
        p->Q = Q;
    }
    std::vector<double> d0;   // the spectral u-solve factors of class 0
    // (Cw, Aw, Ainv) per element -> classes
    {
        const int npk = (nu * nu * (nu * nu + 1)) / 2;
        const size_t len = 2 * nq * nq + npk;
        std::vector<int> cls;
        std::vector<double> uniq;
        dedup(E, len, [&](long e, double* b) {
            for (int ix = 0; ix < nq; ++ix)
                for (int iy = 0; iy < nq; ++iy) {
                    b[iy + nq * ix] = pre->Cw[iy + (long)nq * (e + E * ix)];
                    b[nq * nq + iy + nq * ix] = pre->Aw[iy + (long)nq * (e + E * ix)];
                }
            std::memcpy(b + 2 * nq * nq, pre->Ainv_u + (long)npk * e, npk * sizeof(double));
        }, cls, uniq);
        const int nc = (int)(uniq.size() / len);
        p->ncls_ca = nc;
        const int n = nu * nu;
        std::vector<double> cw((size_t)nc * nq * nq), aw((size_t)nc * nq * nq), ai((size_t)nc * n * n);
        std::vector<double> dfac((size_t)nc * n);
        bool separable = !p->mass_cg;   // conjugate gradients run with the general kernel only
        for (int c = 0; c < nc; ++c) {
            std::memcpy(&cw[(size_t)c * nq * nq], &uniq[(size_t)c * len], nq * nq * sizeof(double));
            std::memcpy(&aw[(size_t)c * nq * nq], &uniq[(size_t)c * len + nq * nq], nq * nq * sizeof(double));
            // unpack the upper triangle (column jj at jj (jj + 1) / 2) into the full symmetric matrix
            const double* pk = &uniq[(size_t)c * len + 2 * nq * nq];
            double* full = &ai[(size_t)c * n * n];
            for (int jj = 0; jj < n; ++jj)
                for (int i = 0; i <= jj; ++i) {
                    const double v = pk[(jj * (jj + 1)) / 2 + i];
                    full[i + n * jj] = v;
                    full[jj + n * i] = v;
                }
            // c^2 constant on the element: cw = c^2 w (x) w; then the inverse has
            // the spectral form (Q (x) Q) diag(D) (Q (x) Q)^T, checked against the
            // stored inverse entry by entry
            if (!separable) continue;
            const double* cwc = &cw[(size_t)c * nq * nq];
            double c2 = 0.0;
            for (int ix = 0; ix < nq; ++ix)
                for (int iy = 0; iy < nq; ++iy) c2 += cwc[iy + nq * ix] / (pre->w[iy] * pre->w[ix]);
            c2 /= nq * nq;
            for (int ix = 0; ix < nq && separable; ++ix)
                for (int iy = 0; iy < nq; ++iy)
                    if (!(std::fabs(cwc[iy + nq * ix] - c2 * pre->w[iy] * pre->w[ix]) <= 1e-13 * std::fabs(cwc[iy + nq * ix])) ||
                        !(c2 > 0.0)) {
                        separable = false;
                        break;
                    }
            if (!separable) continue;
            double* D = &dfac[(size_t)c * n];
            for (int jj = 0; jj < nu; ++jj)
                for (int i = 0; i < nu; ++i)
                    D[i + nu * jj] = 1.0 / (c2 * (p->lam[i] + p->lam[jj]) + (i == 0 && jj == 0 ? 1.0 : 0.0));
            double amax = 0.0, dmax = 0.0;
            for (int col = 0; col < n && separable; ++col) {
                const int a2 = col % nu, b2 = col / nu;
                for (int row = 0; row < n; ++row) {
                    const int a1 = row % nu, b1 = row / nu;
                    double s = 0.0;
                    for (int jj = 0; jj < nu; ++jj)
                        for (int i = 0; i < nu; ++i)
                            s += p->Q[a1 + nu * i] * p->Q[b1 + nu * jj] * D[i + nu * jj] * p->Q[a2 + nu * i] *
                                 p->Q[b2 + nu * jj];
                    const double ref = full[row + n * col];
                    amax = std::max(amax, std::fabs(ref));
                    dmax = std::max(dmax, std::fabs(s - ref));
                }
            }
            if (!(dmax <= 1e-10 * amax)) separable = false;
        }
        const bool fd = separable;
        if (fd) d0.assign(dfac.begin(), dfac.begin() + n);
        UP(p->Cw, cw.data(), cw.size());
        UP(p->Aw, aw.data(), aw.size());
        if (fd) UP(p->Ainv, dfac.data(), dfac.size());
        else UP(p->Ainv, ai.data(), ai.size());
        UP(p->cls_ca, cls.data(), cls.size());
        p->mode = fd ? MODE_SEPARABLE : (p->mass_cg ? MODE_GENERAL_CG : MODE_GENERAL);
    }
    // beta at the high-order points per element -> classes
    {
        const size_t len = (size_t)nh * nh;
        std::vector<int> cls;
        std::vector<double> uniq;
        dedup(E, len, [&](long e, double* b) {
            for (int q = 0; q < nh; ++q)
                for (int pp = 0; pp < nh; ++pp) b[pp + nh * q] = pre->beta_hq[pp + (long)nh * (e + E * q)];
        }, cls, uniq);
        p->ncls_b = (int)(uniq.size() / len);
        UP(p->beta, uniq.data(), uniq.size());
        UP(p->cls_b, cls.data(), cls.size());
    }
    // face arrays: [i][face] with the face index fastest, or one constant
    {
        const double* src[F_COUNT] = {pre->hc2, pre->haf, pre->hcf, pre->hbf, pre->vc2, pre->vaf, pre->vcf,
                                      pre->vbf, pre->hc2m, pre->hafm, pre->vc2m, pre->vafm, pre->hJC, pre->vJC};
        const size_t nf = (size_t)N * (N + 1);
        for (int a = 0; a < F_COUNT; ++a) {
            const double* s = src[a];
            const size_t n = nf * nq;
            if (all_equal(s, n)) {
                double v = s[0];
                UP(p->f[a].p, &v, 1);
                p->f[a].uni = 1;
                p->fu[a] = v;
            } else {
                std::vector<double> t(n);
                for (size_t face = 0; face < nf; ++face)
                    for (int i = 0; i < nq; ++i) t[(size_t)i * nf + face] = s[i + (size_t)nq * face];
                UP(p->f[a].p, t.data(), n);
                p->f[a].uni = 0;
            }
        }
        bool faces_uni = !pre->one_sided;
        for (int a = 0; a < F_COUNT; ++a) faces_uni = faces_uni && p->f[a].uni;
        if (p->mode == MODE_SEPARABLE && p->ncls_ca == 1 && p->ncls_b == 1 && faces_uni) p->mode = MODE_UNIFORM;
    }
    // the split tables of the two-lane kernels; a one-class problem's records
    // as uploaded (Cw, Aw: nq x nq, [iy + nq ix]; beta: nh x nh, [pp + nh q])
    {
        SplitSrc src{pre->w, pre->W2d_h, pre->P_u, pre->S_u, pre->P_h_v, pre->Ebl_u, pre->Ebr_u, pre->E1bl_u,
                     pre->E1br_u, nullptr, nullptr, nullptr, d0.empty() ? nullptr : d0.data()};
        std::vector<double> rc((size_t)nq * nq), ra((size_t)nq * nq), rb((size_t)nh * nh);
        if (p->mode == MODE_UNIFORM) {
            for (int ix = 0; ix < nq; ++ix)
                for (int iy = 0; iy < nq; ++iy) {
                    rc[iy + (size_t)nq * ix] = pre->Cw[iy + (long)nq * (0 + E * ix)];
                    ra[iy + (size_t)nq * ix] = pre->Aw[iy + (long)nq * (0 + E * ix)];
                }
            for (int q = 0; q < nh; ++q)
                for (int pp = 0; pp < nh; ++pp) rb[pp + (size_t)nh * q] = pre->beta_hq[pp + (long)nh * (0 + E * q)];
            src.cw = rc.data(); src.aw = ra.data(); src.beta = rb.data();
        }
        if (build_split_tables(p, src) != hipSuccess) { wedg_problem_destroy(p); return WEDG_ERR_HIP; }
    }
    // boundary data (NQ, N), unchanged
    {
        const double* src[8] = {pre->bcB_sp, pre->bcB_spy, pre->bcT_sp, pre->bcT_spy,
                                pre->bcL_sp, pre->bcL_spx, pre->bcR_sp, pre->bcR_spx};
        for (int a = 0; a < 8; ++a) UP(p->bcd[a], src[a], (size_t)nq * N);
    }
    // projections used by the mode, element-fastest [m][E]; the others enter the
    // reference's sum with an exactly zero factor (src/rhs.jl:344-348)
    {
        const double* src[5] = {pre->PiSP, pre->PiBSP2, pre->PiCdiv, pre->PiDdiv, pre->src_projE};
        int use[5] = {0, 0, 0, 0, 0};
        if (pre->mms_on) { use[0] = use[1] = use[2] = use[3] = 1; }
        else use[2] = 1;                 // (1 - ct) with ct = 0
        if (pre->src_on) use[4] = 1;
        const size_t n2 = (size_t)nv * nv;
        for (int a = 0; a < 5; ++a) {
            if (!use[a]) continue;
            std::vector<double> t(n2 * E);
            bool zero = true;
            for (long e = 0; e < E; ++e)
                for (size_t m = 0; m < n2; ++m) {
                    double v = src[a][m + n2 * e];
                    t[m * E + e] = v;
                    if (v != 0.0) zero = false;
                }
            if (zero && !(pre->mms_on && a < 4)) continue;
            UP(p->proj[a], t.data(), t.size());
            p->has_proj[a] = 1;
        }
    }
#undef UP
    (void)hipHostFree(p->stage);
    p->stage = nullptr;
    *out = p;
    return WEDG_OK;
}

extern "C" wedg_status wedg_problem_info(const wedg_problem* p, int* mode, int* classes, int* two_lane) {
    if (!p) return WEDG_ERR_ARGUMENT;
    if (mode) *mode = p->mode;
    if (classes) *classes = p->ncls_ca;
    if (two_lane) *two_lane = p->tab2 && !p->mass_cg && (p->mode == MODE_UNIFORM || p->mode == MODE_SEPARABLE);
    return WEDG_OK;
}

extern "C" wedg_status wedg_problem_destroy(wedg_problem* p) {
    if (!p) return WEDG_OK;
    for (void* a : p->allocs) (void)hipFree(a);
    if (p->stage) (void)hipHostFree(p->stage);
    delete p;
    return WEDG_OK;
}

wedg_status wedg::reset_err(wedg_workspace* ws) {
    // the pinned record may still be the source of a copy in flight
    HIPOK(hipStreamSynchronize(ws->stream));
    DevErr& h = *ws->herr;
    std::memset(&h, 0, sizeof(h));
    h.fkey = ~0ull;
    h.nonfinite = INT_MAX;
    h.nf_serial = INT_MAX;
    ws->stages = 0;
    ws->serial = 0;   // k_prep forms the count slot anew at the start of every call
    HIPOK(hipMemcpyAsync(ws->err, &h, sizeof(h), hipMemcpyHostToDevice, ws->stream));
    HIPOK(hipMemsetAsync(ws->wstat, 0, 2 * ws->nwaves_cap() * sizeof(unsigned), ws->stream));
    return WEDG_OK;
}

// host <-> device copies of user arrays through the pinned staging buffers
static wedg_status stage_in(wedg_workspace* ws, double* d, double* pinned, const double* h, size_t n) {
    HIPOK(hipStreamSynchronize(ws->stream));
    std::memcpy(pinned, h, n * sizeof(double));
    HIPOK(hipMemcpyAsync(d, pinned, n * sizeof(double), hipMemcpyHostToDevice, ws->stream));
    return WEDG_OK;
}

static wedg_status stage_out(wedg_workspace* ws, double* h, double* pinned, const double* d, size_t n) {
    HIPOK(hipMemcpyAsync(pinned, d, n * sizeof(double), hipMemcpyDeviceToHost, ws->stream));
    HIPOK(hipStreamSynchronize(ws->stream));
    std::memcpy(h, pinned, n * sizeof(double));
    return WEDG_OK;
}

extern "C" wedg_status wedg_workspace_create(wedg_problem* p, int cheb_policy, wedg_workspace** out) {
    if (!p || !out) return WEDG_ERR_ARGUMENT;
    if (cheb_policy != WEDG_CHEB_GLOBAL && cheb_policy != WEDG_CHEB_ELEMENT) return WEDG_ERR_ARGUMENT;
    auto* ws = new wedg_workspace();
    ws->p = p;
    ws->cheb_global = cheb_policy == WEDG_CHEB_GLOBAL;
    hipError_t he = hipStreamCreateWithFlags(&ws->stream, hipStreamNonBlocking);
    if (he != hipSuccess) { delete ws; return WEDG_ERR_HIP; }
    auto A = [&](size_t n) -> double* {
        double* d = dalloc<double>(n, &he);
        if (he == hipSuccess) { ws->allocs.push_back(d); (void)hipMemset(d, 0, n * sizeof(double)); }
        return d;
    };
    const size_t nu = ws->nU(), nv = ws->nV();
    ws->yu = A(nu); ws->yv = A(nv); ws->su = A(nu); ws->sv = A(nv);
    ws->au = A(nu); ws->av = A(nv); ws->ou = A(nu); ws->ov = A(nv);
    ws->bsl = A((size_t)4 * p->nu * 3 * p->N);
    if (he == hipSuccess) {
        void* e = nullptr;
        he = hipMalloc(&e, sizeof(DevErr) + 16);
        if (he == hipSuccess) { ws->allocs.push_back(e); ws->err = (DevErr*)e; }
    }
    if (he == hipSuccess) he = hipHostMalloc((void**)&ws->herr, sizeof(DevErr), hipHostMallocDefault);
    if (he == hipSuccess) he = hipHostMalloc((void**)&ws->hu, nu * sizeof(double), hipHostMallocDefault);
    if (he == hipSuccess) he = hipHostMalloc((void**)&ws->hv, nv * sizeof(double), hipHostMallocDefault);
    if (he == hipSuccess) {
        void* c = nullptr;
        const size_t nw = ws->nwaves_cap();   // slots for either stage kernel's waves
        he = hipMalloc(&c, 2 * nw * sizeof(int) + 2 * nw * sizeof(unsigned));
        if (he == hipSuccess) {
            ws->allocs.push_back(c);
            ws->cnt = (int*)c;
            ws->wstat = (unsigned*)((int*)c + 2 * nw);
            he = hipMemset(c, 0, 2 * nw * sizeof(int) + 2 * nw * sizeof(unsigned));
        }
        if (he == hipSuccess) he = hipHostMalloc((void**)&ws->hstat, 2 * nw * sizeof(unsigned), hipHostMallocDefault);
    }
    // the fills above run on the null stream; finish them before the workspace stream is used
    if (he == hipSuccess) he = hipDeviceSynchronize();
    if (he != hipSuccess) { wedg_workspace_destroy(ws); return WEDG_ERR_HIP; }
    if (reset_err(ws) != WEDG_OK || hipStreamSynchronize(ws->stream) != hipSuccess) {
        wedg_workspace_destroy(ws);
        return WEDG_ERR_HIP;
    }
    *out = ws;
    return WEDG_OK;
}

static void graph_free(wedg_workspace* ws);

extern "C" wedg_status wedg_workspace_destroy(wedg_workspace* ws) {
    if (!ws) return WEDG_OK;
    if (ws->stream) (void)hipStreamSynchronize(ws->stream);
    graph_free(ws);
    for (auto& u : ws->ev_used) { (void)hipEventDestroy(u.second.first); (void)hipEventDestroy(u.second.second); }
    for (auto& e : ws->ev_pool) (void)hipEventDestroy(e);
    for (void* a : ws->allocs) (void)hipFree(a);
    if (ws->herr) (void)hipHostFree(ws->herr);
    if (ws->hu) (void)hipHostFree(ws->hu);
    if (ws->hv) (void)hipHostFree(ws->hv);
    if (ws->hstat) (void)hipHostFree(ws->hstat);
    if (ws->hrecv) (void)hipHostFree(ws->hrecv);
    if (ws->hdiag) (void)hipHostFree(ws->hdiag);
    if (ws->hrows) (void)hipHostFree(ws->hrows);
    if (ws->hnorm) (void)hipHostFree(ws->hnorm);
    if (ws->drecv) (void)hipFree(ws->drecv);
    if (ws->stream) (void)hipStreamDestroy(ws->stream);
    delete ws;
    return WEDG_OK;
}

KArgs wedg::problem_args(const wedg_problem* p, int cheb_global, double* bsl, DevErr* err, int* cnt,
                         unsigned* wstat) {
    KArgs a;
    std::memset(&a, 0, sizeof(a));
    a.N = p->N; a.E = p->E;
    a.cheb_global = p->mass_cg ? 0 : cheb_global;   // no counts with conjugate gradients
    a.mass_cg = p->mass_cg;
    a.mass_tol = p->mass_tol;
    a.one_sided = p->one_sided;
    const double h = 1.0 / p->N;
    a.two_h = 2.0 / h; a.four_hh = 4.0 / (h * h);
    a.theta = p->theta; a.tau = p->tau; a.lu = p->lu; a.lv = p->lv; a.omega = p->omega;
    a.cheb_L = std::acosh(1.0 / p->mass_tol);
    for (int s = 0; s < 4; ++s) {
        a.kind[s] = p->bc[s].kind; a.gam[s] = p->bc[s].gamma_abc; a.eta[s] = p->bc[s].eta_abc;
        a.sig[s] = p->bc[s].sigma; a.kb[s] = p->bc[s].kb; a.kap[s] = p->bc[s].kappa;
    }
    a.tab = p->tab;
    a.tab2 = p->tab2;
    a.cls_ca = p->cls_ca; a.Cw = p->Cw; a.Aw = p->Aw; a.Ainv = p->Ainv;
    a.cls_b = p->cls_b; a.beta = p->beta;
    for (int i = 0; i < F_COUNT; ++i) { a.f[i] = p->f[i]; a.fu[i] = p->fu[i]; }
    a.fstride = p->N * (p->N + 1);
    for (int i = 0; i < 8; ++i) a.bcd[i] = p->bcd[i];
    a.bsl = bsl; a.err = err; a.cnt = cnt; a.wstat = wstat;
    return a;
}

KArgs wedg::base_args(wedg_workspace* ws) {
    return problem_args(ws->p, ws->cheb_global, ws->bsl, ws->err, ws->cnt, ws->wstat);
}

// Time factors of the stage at time t and the projection coefficients, in the
// order of the reference's sum (src/rhs.jl:344-348, 509-511).
void wedg::stage_time(const wedg_problem* p, double t, KArgs& a) {
    a.ct = p->mms_on ? std::cos(p->omega * t) : 0.0;
    a.st = p->mms_on ? std::sin(p->omega * t) : 0.0;
    double srct = 0.0;
    if (p->src_on && t < p->src_Ts) {
        const double sv = std::sin(M_PI * t / p->src_Ts);
        srct = std::sin(p->src_omega * t) * (sv * sv);
    }
    a.srct = srct;
    const double om = p->omega, ct = a.ct, st = a.st;
    const double coef[5] = {-(om * om) * ct, -2.0 * (om * om * om) * st * ct, 1.0 - ct, om * st, srct};
    // A projection whose time factor is zero is left out: its term is a signed
    // zero, which changes no sum it is added to except possibly the sign of a
    // zero (in :initial_value mode three of the four manufactured-solution
    // projections, 864 B per element and stage at q_v = 5).
    a.nproj = 0;
    for (int i = 0; i < 5; ++i)
        if (p->has_proj[i] && coef[i] != 0.0) {
            a.pc[a.nproj] = coef[i];
            a.proj[a.nproj] = p->proj[i];
            ++a.nproj;
        }
}

static hipEvent_t ev_get(wedg_workspace* ws) {
    if (!ws->ev_pool.empty()) {
        hipEvent_t e = ws->ev_pool.back();
        ws->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
}

wedg_status wedg::run_kernel(wedg_workspace* ws, const KArgs& a, int which) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ws->timing) {
        e0 = ev_get(ws);
        e1 = ev_get(ws);
        HIPOK(hipEventRecord(e0, ws->stream));
    }
    HIPOK(launch(a, ws->p, which, ws->stream, ws->two()));
    if (ws->timing) {
        HIPOK(hipEventRecord(e1, ws->stream));
        ws->ev_used.push_back({which, {e0, e1}});
    }
    return WEDG_OK;
}

// The arguments of one stage: k_stage on Y_s, writing the stage output and,
// when has_next, preparing the stage that follows at time t_next. The output
// must not alias Y_s: a stage reads its neighbours' input state.
KArgs wedg::stage_args(wedg_workspace* ws, int stage, const double* yu, const double* yv,
                        const double* y0u, const double* y0v, double* outu, double* outv,
                        double t, double c, long step, bool has_next, double t_next) {
    KArgs a = base_args(ws);
    a.stage = stage;
    a.serial = ++ws->serial;
    if (ws->serial > 0x3fffffff) ws->serial = 1;
    a.slot = a.serial & 1;
    a.step = (int)step;
    a.cstage = c;
    stage_time(ws->p, t, a);
    a.yu = yu; a.yv = yv; a.y0u = y0u; a.y0v = y0v;
    a.accu = ws->au; a.accv = ws->av; a.outu = outu; a.outv = outv;
    a.has_next = has_next ? 1 : 0;
    if (has_next) {
        KArgs nx;
        std::memset(&nx, 0, sizeof(nx));
        stage_time(ws->p, t_next, nx);
        a.nct = nx.ct;
        a.nst = nx.st;
    }
    ++ws->stages;
    return a;
}

// One stage (after k_prep when it is the first stage of a call).
static wedg_status run_stage(wedg_workspace* ws, int stage, const double* yu, const double* yv,
                             const double* y0u, const double* y0v, double* outu, double* outv,
                             double t, double c, long step, bool first, bool has_next, double t_next) {
    const KArgs a = stage_args(ws, stage, yu, yv, y0u, y0v, outu, outv, t, c, step, has_next, t_next);
    wedg_status st;
    if (first && (st = run_kernel(ws, a, 0)) != WEDG_OK) return st;
    return run_kernel(ws, a, 1);
}

// Stage s (1..4) of one RK4 step (src/solver.jl:300-321) on the device state:
// the stage values alternate between the S (su, sv) and O (ou, ov) buffers, so
// no stage writes the state it reads; the last stage updates y in place (it
// reads only its own element of y). last: the last step of a call.
static KArgs rk4_args(wedg_workspace* ws, int s, double t, double dt, long step, bool last) {
    const double hd = 0.5 * dt, th = t + 0.5 * dt, tf = t + dt;
    switch (s) {
        case 1: return stage_args(ws, 1, ws->yu, ws->yv, ws->yu, ws->yv, ws->su, ws->sv, t, hd, step, true, th);
        case 2: return stage_args(ws, 2, ws->su, ws->sv, ws->yu, ws->yv, ws->ou, ws->ov, th, hd, step, true, th);
        case 3: return stage_args(ws, 3, ws->ou, ws->ov, ws->yu, ws->yv, ws->su, ws->sv, th, dt, step, true, tf);
        default: return stage_args(ws, 4, ws->su, ws->sv, ws->yu, ws->yv, ws->yu, ws->yv, tf, dt / 6.0, step, !last, tf);
    }
}

// One RK4 step; first: the first step of a call.
wedg_status wedg::rk4_device(wedg_workspace* ws, double t, double dt, long step, bool first, bool last) {
    wedg_status st;
    for (int s = 1; s <= 4; ++s) {
        const KArgs a = rk4_args(ws, s, t, dt, step, last);
        if (s == 1 && first && (st = run_kernel(ws, a, 0)) != WEDG_OK) return st;
        if ((st = run_kernel(ws, a, 1)) != WEDG_OK) return st;
    }
    return WEDG_OK;
}

// One RK4 step from the state (inu, inv) into (outu, outv), the same stages as
// rk4_device with the input and output anywhere (the output may be the input:
// the last stage reads only its own element of it).
wedg_status wedg::rk4_step_io(wedg_workspace* ws, const double* inu, const double* inv, double* outu, double* outv,
                              double t, double dt, long step, bool first, bool last) {
    const double hd = 0.5 * dt, th = t + 0.5 * dt, tf = t + dt;
    wedg_status st;
    for (int s = 1; s <= 4; ++s) {
        KArgs a;
        switch (s) {
            case 1: a = stage_args(ws, 1, inu, inv, inu, inv, ws->su, ws->sv, t, hd, step, true, th); break;
            case 2: a = stage_args(ws, 2, ws->su, ws->sv, inu, inv, ws->ou, ws->ov, th, hd, step, true, th); break;
            case 3: a = stage_args(ws, 3, ws->ou, ws->ov, inu, inv, ws->su, ws->sv, th, dt, step, true, tf); break;
            default: a = stage_args(ws, 4, ws->su, ws->sv, inu, inv, outu, outv, tf, dt / 6.0, step, !last, tf);
        }
        if (s == 1 && first && (st = run_kernel(ws, a, 0)) != WEDG_OK) return st;
        if ((st = run_kernel(ws, a, 1)) != WEDG_OK) return st;
    }
    return WEDG_OK;
}

// Reads the failure record and maps it to a status (boundary before mass,
// src/rhs.jl:530-531).
wedg_status wedg::collect(wedg_workspace* ws, int nonfinite_is_error) {
    HIPOK(hipStreamSynchronize(ws->stream));
    if (ws->timing) {
        for (auto& u : ws->ev_used) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, u.second.first, u.second.second) == hipSuccess) {
                ws->t_ms[u.first] += ms;
                ws->t_n[u.first] += 1;
            }
            ws->ev_pool.push_back(u.second.first);
            ws->ev_pool.push_back(u.second.second);
        }
        ws->ev_used.clear();
    }
    HIPOK(hipMemcpyAsync(ws->herr, ws->err, sizeof(DevErr), hipMemcpyDeviceToHost, ws->stream));
    HIPOK(hipStreamSynchronize(ws->stream));
    const DevErr h = *ws->herr;
    wedg_error_info& L = ws->last;
    std::memset(&L, 0, sizeof(L));
    {
        const size_t nw = ws->nwaves();
        HIPOK(hipMemcpyAsync(ws->hstat, ws->wstat, 2 * nw * sizeof(unsigned), hipMemcpyDeviceToHost, ws->stream));
        HIPOK(hipStreamSynchronize(ws->stream));
        unsigned long long sum = 0;
        unsigned mx = 0;
        for (size_t w = 0; w < nw; ++w) {
            sum += ws->hstat[w];
            mx = std::max(mx, ws->hstat[nw + w]);
        }
        L.cheb_steps = (int)mx;
        L.cg_unconverged = h.cg_unconverged;
        ws->cheb_wave_sum = sum;
    }
    return (wedg_status)map_failure(h, ws->p->nu, ws->p->N, nonfinite_is_error, L);
}

extern "C" wedg_status wedg_rhs_device(wedg_workspace* ws, const double* u, const double* v, double t,
                                       double* ut, double* vt) {
    if (!ws || !u || !v || !ut || !vt) return WEDG_ERR_ARGUMENT;
    wedg_status s;
    if ((s = reset_err(ws))) return s;
    return run_stage(ws, 0, u, v, u, v, ut, vt, t, 0.0, 0, true, false, t);
}

extern "C" wedg_status wedg_sync(wedg_workspace* ws) {
    if (!ws) return WEDG_ERR_ARGUMENT;
    return collect(ws, 0);
}

extern "C" wedg_status wedg_rhs(wedg_workspace* ws, const double* u, const double* v, double t, int N,
                                double h, double* ut, double* vt) {
    if (!ws || !u || !v || !ut || !vt) return WEDG_ERR_ARGUMENT;
    if (N != ws->p->N || h != 1.0 / N) return WEDG_ERR_DIMENSION;
    wedg_status s;
    if ((s = stage_in(ws, ws->su, ws->hu, u, ws->nU()))) return s;
    if ((s = stage_in(ws, ws->sv, ws->hv, v, ws->nV()))) return s;
    if ((s = reset_err(ws))) return s;
    if ((s = run_stage(ws, 0, ws->su, ws->sv, ws->su, ws->sv, ws->ou, ws->ov, t, 0.0, 0, true, false, t))) return s;
    if ((s = collect(ws, 0))) return s;
    if ((s = stage_out(ws, ut, ws->hu, ws->ou, ws->nU()))) return s;
    return stage_out(ws, vt, ws->hv, ws->ov, ws->nV());
}

extern "C" wedg_status wedg_set_state(wedg_workspace* ws, const double* u, const double* v) {
    if (!ws || !u || !v) return WEDG_ERR_ARGUMENT;
    wedg_status s;
    if ((s = stage_in(ws, ws->yu, ws->hu, u, ws->nU()))) return s;
    if ((s = stage_in(ws, ws->yv, ws->hv, v, ws->nV()))) return s;
    HIPOK(hipStreamSynchronize(ws->stream));
    return WEDG_OK;
}

extern "C" wedg_status wedg_get_state(wedg_workspace* ws, double* u, double* v) {
    if (!ws || !u || !v) return WEDG_ERR_ARGUMENT;
    wedg_status s;
    if ((s = stage_out(ws, u, ws->hu, ws->yu, ws->nU()))) return s;
    return stage_out(ws, v, ws->hv, ws->yv, ws->nV());
}

extern "C" wedg_status wedg_state_device(wedg_workspace* ws, double** u, double** v) {
    if (!ws || !u || !v) return WEDG_ERR_ARGUMENT;
    *u = ws->yu;
    *v = ws->yv;
    return WEDG_OK;
}

static void graph_free(wedg_workspace* ws) {
    for (int r = 0; r < wedg_workspace::kGraphRing; ++r) {
        if (ws->gused[r]) (void)hipEventSynchronize(ws->gdone[r]);
        if (ws->gexec[r]) (void)hipGraphExecDestroy(ws->gexec[r]);
        if (ws->gdone[r]) (void)hipEventDestroy(ws->gdone[r]);
        ws->gexec[r] = nullptr;
        ws->gdone[r] = nullptr;
        ws->gused[r] = false;
    }
    if (ws->graph) (void)hipGraphDestroy(ws->graph);
    ws->graph = nullptr;
    ws->graph_p = nullptr;
    ws->gnodes.clear();
    ws->gparams.clear();
    ws->glaunch = 0;
}

// Captures the stage launches of args (one k_stage each) into the workspace's
// graph: the kernel nodes in launch order and the executable copies.
static wedg_status graph_capture(wedg_workspace* ws, const std::vector<KArgs>& args) {
    HIPOK(hipStreamBeginCapture(ws->stream, hipStreamCaptureModeThreadLocal));
    hipError_t le = hipSuccess;
    for (const KArgs& a : args)
        if (le == hipSuccess) le = launch(a, ws->p, 1, ws->stream, ws->two());
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(ws->stream, &g);
    if (le != hipSuccess || ce != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        return WEDG_ERR_HIP;
    }
    ws->graph = g;
    ws->graph_p = ws->p;
    // a chain (one stream): the root, then each node's one dependent
    size_t nr = 0;
    HIPOK(hipGraphGetRootNodes(g, nullptr, &nr));
    if (nr != 1) return WEDG_ERR_HIP;
    hipGraphNode_t node;
    HIPOK(hipGraphGetRootNodes(g, &node, &nr));
    for (;;) {
        hipKernelNodeParams kp;
        HIPOK(hipGraphKernelNodeGetParams(node, &kp));
        ws->gnodes.push_back(node);
        ws->gparams.push_back(kp);
        size_t nd = 0;
        HIPOK(hipGraphNodeGetDependentNodes(node, nullptr, &nd));
        if (nd == 0) break;
        if (nd != 1) return WEDG_ERR_HIP;
        HIPOK(hipGraphNodeGetDependentNodes(node, &node, &nd));
    }
    if (ws->gnodes.size() != args.size()) return WEDG_ERR_HIP;
    for (int r = 0; r < wedg_workspace::kGraphRing; ++r) {
        HIPOK(hipGraphInstantiate(&ws->gexec[r], g, nullptr, nullptr, 0));
        HIPOK(hipEventCreateWithFlags(&ws->gdone[r], hipEventDisableTiming));
    }
    return WEDG_OK;
}

// The stage launches of args through the graph: capture on first use (the
// captured launches do not run), then each launch sets the arguments of an
// executable copy whose previous launch has ended and launches it.
static wedg_status graph_run(wedg_workspace* ws, const std::vector<KArgs>& args) {
    if (ws->graph && (ws->graph_p != ws->p || ws->gnodes.size() != args.size())) graph_free(ws);
    if (!ws->graph) {
        const wedg_status st = graph_capture(ws, args);
        if (st) {
            graph_free(ws);
            return st;
        }
    }
    const int r = (int)(ws->glaunch % wedg_workspace::kGraphRing);
    if (ws->gused[r]) HIPOK(hipEventSynchronize(ws->gdone[r]));
    for (size_t i = 0; i < args.size(); ++i) {
        hipKernelNodeParams kp = ws->gparams[i];
        void* kargs_ptr[1] = {const_cast<KArgs*>(&args[i])};
        kp.kernelParams = kargs_ptr;
        kp.extra = nullptr;
        HIPOK(hipGraphExecKernelNodeSetParams(ws->gexec[r], ws->gnodes[i], &kp));
    }
    HIPOK(hipGraphLaunch(ws->gexec[r], ws->stream));
    HIPOK(hipEventRecord(ws->gdone[r], ws->stream));
    ws->gused[r] = true;
    ++ws->glaunch;
    return WEDG_OK;
}

extern "C" wedg_status wedg_workspace_set_lanes_per_element(wedg_workspace* ws, int lanes) {
    if (!ws || (lanes != 1 && lanes != 2)) return WEDG_ERR_ARGUMENT;
    if (lanes != ws->lpe) {
        HIPOK(hipStreamSynchronize(ws->stream));
        graph_free(ws);   // a captured graph holds the other kernels
        ws->lpe = lanes;
    }
    return WEDG_OK;
}

extern "C" wedg_status wedg_workspace_lanes_per_element(const wedg_workspace* ws, int* lanes) {
    if (!ws || !lanes) return WEDG_ERR_ARGUMENT;
    *lanes = ws->two() ? 2 : 1;
    return WEDG_OK;
}

extern "C" wedg_status wedg_graph_enable(wedg_workspace* ws, long chunk_steps) {
    if (!ws || chunk_steps < 0 || chunk_steps > 4096) return WEDG_ERR_ARGUMENT;
    if (chunk_steps != ws->graph_steps) {
        HIPOK(hipStreamSynchronize(ws->stream));
        graph_free(ws);
    }
    ws->graph_steps = chunk_steps;
    return WEDG_OK;
}

extern "C" wedg_status wedg_advance_async(wedg_workspace* ws, double t0, double dt, long nsteps,
                                          int stop_on_nonfinite, double* t_end) {
    if (!ws || nsteps < 0) return WEDG_ERR_ARGUMENT;
    (void)stop_on_nonfinite;
    wedg_status s;
    if ((s = reset_err(ws))) return s;
    double t = t0;
    long it = 0;
    // whole chunks of graph_steps steps through the graph (the same arguments,
    // in the same order, as the launches below)
    const long K = ws->graph_steps;
    if (K > 0 && !ws->timing) {
        std::vector<KArgs> args((size_t)4 * K);
        while (nsteps - it >= K) {
            KArgs prep{};
            const bool first = it == 0;
            for (long k = 0; k < K; ++k, ++it) {
                for (int st = 1; st <= 4; ++st) {
                    args[(size_t)4 * k + st - 1] = rk4_args(ws, st, t, dt, it + 1, it == nsteps - 1);
                    if (first && k == 0 && st == 1) prep = args[0];
                }
                t += dt;
            }
            if (first) HIPOK(launch(prep, ws->p, 0, ws->stream, ws->two()));
            if ((s = graph_run(ws, args))) return s;
        }
    }
    for (; it < nsteps; ++it) {
        if ((s = rk4_device(ws, t, dt, it + 1, it == 0, it == nsteps - 1))) return s;
        t += dt;
    }
    if (t_end) *t_end = t;
    return WEDG_OK;
}

extern "C" wedg_status wedg_advance(wedg_workspace* ws, double t0, double dt, long nsteps, int stop_on_nonfinite,
                                    double* t_end) {
    wedg_status s = wedg_advance_async(ws, t0, dt, nsteps, stop_on_nonfinite, t_end);
    if (s) return s;
    return collect(ws, stop_on_nonfinite);
}

extern "C" wedg_status wedg_rk4_step(wedg_workspace* ws, double* u, double* v, double t, double dt, int N, double h,
                                     int* finite) {
    if (!ws || !u || !v) return WEDG_ERR_ARGUMENT;
    if (N != ws->p->N || h != 1.0 / N) return WEDG_ERR_DIMENSION;
    wedg_status s;
    if ((s = wedg_set_state(ws, u, v))) return s;
    if ((s = reset_err(ws))) return s;
    if ((s = rk4_device(ws, t, dt, 1, true, true))) return s;
    s = collect(ws, 0);
    if (finite) *finite = ws->last.step == 0;
    if (s) return s;
    return wedg_get_state(ws, u, v);
}

extern "C" wedg_status wedg_last_error(const wedg_workspace* ws, wedg_error_info* info) {
    if (!ws || !info) return WEDG_ERR_ARGUMENT;
    *info = ws->last;
    return WEDG_OK;
}

extern "C" wedg_status wedg_timing_enable(wedg_workspace* ws, int on) {
    if (!ws) return WEDG_ERR_ARGUMENT;
    ws->timing = on;
    for (int i = 0; i < 4; ++i) { ws->t_ms[i] = 0; ws->t_n[i] = 0; }
    return WEDG_OK;
}

extern "C" wedg_status wedg_timing_read(wedg_workspace* ws, int family, double* ms, long* launches) {
    if (!ws || family < 0 || family > 3) return WEDG_ERR_ARGUMENT;
    if (ms) *ms = ws->t_ms[family];
    if (launches) *launches = ws->t_n[family];
    return WEDG_OK;
}

extern "C" wedg_status wedg_cheb_stats(const wedg_workspace* ws, double* mean_steps, long* stages) {
    if (!ws) return WEDG_ERR_ARGUMENT;
    const long waves = (long)ws->nwaves();
    if (stages) *stages = ws->stages;
    if (mean_steps)
        *mean_steps = ws->stages > 0 ? (double)ws->cheb_wave_sum / ((double)waves * ws->stages) : 0.0;
    return WEDG_OK;
}

namespace {
__global__ void k_copy_probe(const double* __restrict__ src, double* __restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
}  // namespace

extern "C" wedg_status wedg_probe_copy(size_t n, int repeats, double* ms) {
    double *a = nullptr, *b = nullptr;
    HIPOK(hipMalloc(&a, n * sizeof(double)));
    if (hipMalloc(&b, n * sizeof(double)) != hipSuccess) { (void)hipFree(a); return WEDG_ERR_HIP; }
    hipEvent_t e0, e1;
    wedg_status st = WEDG_OK;
    if (hipMemset(a, 0, n * sizeof(double)) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        st = WEDG_ERR_HIP;
    } else {
        k_copy_probe<<<1024, 256>>>(a, b, n);
        (void)hipEventRecord(e0, nullptr);
        for (int r = 0; r < repeats; ++r) k_copy_probe<<<1024, 256>>>(a, b, n);
        (void)hipEventRecord(e1, nullptr);
        float t = 0.f;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess) st = WEDG_ERR_HIP;
        if (ms) *ms = repeats > 0 ? t / repeats : 0.0;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    return st;
}

extern "C" const char* wedg_version(void) { return "wedg 0.1 gfx950 fp64"; }

extern "C" wedg_status wedg_device_info(char* name, size_t len, int* cu_count, double* hbm_gbs) {
    int dev = 0;
    HIPOK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPOK(hipGetDeviceProperties(&prop, dev));
    if (name && len) {
        std::strncpy(name, prop.gcnArchName, len - 1);
        name[len - 1] = 0;
    }
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (hbm_gbs) *hbm_gbs = 2.0 * prop.memoryClockRate * 1e3 * (prop.memoryBusWidth / 8) / 1e9;
    return WEDG_OK;
}

// ---------------------------------------------------------------------------
// Source-mode diagnostics (src/solver.jl:856-958).

static hipError_t launch_diag(wedg_problem* p, const DiagArgs* d, const RecvArgs* r, const NormArgs* n,
                              hipStream_t s) {
    switch (p->qu) {
        case 1: return launch_diag_q1(d, r, n, p->nv, s);
        case 2: return launch_diag_q2(d, r, n, p->nv, s);
        case 3: return launch_diag_q3(d, r, n, p->nv, s);
        case 4: return launch_diag_q4(d, r, n, p->nv, s);
        case 5: return launch_diag_q5(d, r, n, p->nv, s);
        case 6: return launch_diag_q6(d, r, n, p->nv, s);
        default: return hipErrorInvalidValue;
    }
}

extern "C" wedg_status wedg_problem_set_diag(wedg_problem* p, const double* err_c2, const double* err_beta,
                                             const double* W2d) {
    if (!p || !err_c2 || !err_beta || !W2d) return WEDG_ERR_ARGUMENT;
    const int N = p->N, nq = p->nu;
    const long E = p->E;
    const size_t len = 2 * (size_t)nq * nq;
    std::vector<int> cls;
    std::vector<double> uniq;
    // geo.err_c2, geo.err_beta are (N, N, nq, nq): [j, k, p, q] with j fastest
    dedup(E, len, [&](long e, double* b) {
        const long j = e % N, k = e / N;
        for (int q = 0; q < nq; ++q)
            for (int pp = 0; pp < nq; ++pp) {
                const long idx = j + (long)N * (k + (long)N * (pp + (long)nq * q));
                b[pp + nq * q] = err_c2[idx];
                b[nq * nq + pp + nq * q] = err_beta[idx];
            }
    }, cls, uniq);
    const int nc = (int)(uniq.size() / len);
    std::vector<double> c2((size_t)nc * nq * nq), bq((size_t)nc * nq * nq);
    for (int c = 0; c < nc; ++c) {
        std::memcpy(&c2[(size_t)c * nq * nq], &uniq[(size_t)c * len], nq * nq * sizeof(double));
        std::memcpy(&bq[(size_t)c * nq * nq], &uniq[(size_t)c * len + nq * nq], nq * nq * sizeof(double));
    }
    p->stage_bytes = (size_t)1 << 20;
    if (hipHostMalloc(&p->stage, p->stage_bytes, hipHostMallocDefault) != hipSuccess) return WEDG_ERR_HIP;
    hipError_t he = hipSuccess;
    p->dcls = p->upload(cls.data(), cls.size(), &he);
    if (he == hipSuccess) p->dc2 = p->upload(c2.data(), c2.size(), &he);
    if (he == hipSuccess) p->dbq = p->upload(bq.data(), bq.size(), &he);
    if (he == hipSuccess) p->dW2d = p->upload(W2d, (size_t)nq * nq, &he);
    (void)hipHostFree(p->stage);
    p->stage = nullptr;
    return he == hipSuccess ? WEDG_OK : WEDG_ERR_HIP;
}

// room for `need` receiver values on the device and in pinned memory
static wedg_status recv_capacity(wedg_workspace* ws, size_t need) {
    HIPOK(hipStreamSynchronize(ws->stream));
    if (need > ws->nrecv_cap) {
        if (ws->drecv) (void)hipFree(ws->drecv);
        if (ws->hrecv) (void)hipHostFree(ws->hrecv);
        ws->drecv = nullptr;
        ws->hrecv = nullptr;
        ws->nrecv_cap = 0;
        HIPOK(hipMalloc((void**)&ws->drecv, need * sizeof(double)));
        HIPOK(hipHostMalloc((void**)&ws->hrecv, need * sizeof(double), hipHostMallocDefault));
        ws->nrecv_cap = need;
    }
    return WEDG_OK;
}

// The receivers' elements and basis rows on the device (checked against the mesh).
static wedg_status recv_rows(wedg_workspace* ws, int nrecv, const int* recv_j, const int* recv_k,
                             const double* rows) {
    if (nrecv < 0 || (nrecv > 0 && (!recv_j || !recv_k || !rows))) return WEDG_ERR_ARGUMENT;
    wedg_problem* p = ws->p;
    const int N = p->N, nv = p->nv;
    for (int r = 0; r < nrecv; ++r)
        if (recv_j[r] < 0 || recv_j[r] >= N || recv_k[r] < 0 || recv_k[r] >= N) return WEDG_ERR_ARGUMENT;
    HIPOK(hipStreamSynchronize(ws->stream));
    const size_t rbytes = (size_t)nrecv * (sizeof(int) + 2 * nv * sizeof(double));
    if (nrecv > ws->nrecv_dev) {   // receiver basis rows, then elements, on the device
        if (ws->hrows) (void)hipHostFree(ws->hrows);
        ws->hrows = nullptr;
        ws->nrecv_dev = 0;
        void* d = nullptr;
        HIPOK(hipMalloc(&d, rbytes));
        ws->allocs.push_back(d);
        HIPOK(hipHostMalloc(&ws->hrows, rbytes, hipHostMallocDefault));
        ws->drecv_rows = (double*)d;
        ws->nrecv_dev = nrecv;
    }
    if (nrecv > 0) {
        ws->drecv_elem = (int*)(ws->drecv_rows + (size_t)2 * nv * nrecv);
        double* hr = (double*)ws->hrows;
        int* he_ = (int*)(hr + (size_t)2 * nv * nrecv);
        std::memcpy(hr, rows, (size_t)2 * nv * nrecv * sizeof(double));
        for (int r = 0; r < nrecv; ++r) he_[r] = recv_j[r] + N * recv_k[r];
        HIPOK(hipMemcpyAsync(ws->drecv_rows, ws->hrows, rbytes, hipMemcpyHostToDevice, ws->stream));
        HIPOK(hipStreamSynchronize(ws->stream));
    }
    return WEDG_OK;
}

static RecvArgs recv_args(wedg_workspace* ws, int nrecv) {
    RecvArgs ra;
    ra.v = ws->yv; ra.E = ws->p->E; ra.nrecv = nrecv; ra.elem = ws->drecv_elem; ra.rows = ws->drecv_rows;
    ra.out = ws->drecv; ra.step = 0;
    return ra;
}

static wedg_status recv_out(wedg_workspace* ws, double* recv, size_t need) {
    if (need == 0) return WEDG_OK;
    HIPOK(hipMemcpyAsync(ws->hrecv, ws->drecv, need * sizeof(double), hipMemcpyDeviceToHost, ws->stream));
    HIPOK(hipStreamSynchronize(ws->stream));
    std::memcpy(recv, ws->hrecv, need * sizeof(double));
    return WEDG_OK;
}

extern "C" wedg_status wedg_advance_receivers(wedg_workspace* ws, double t0, double dt, long nsteps, int nrecv,
                                              const int* recv_j, const int* recv_k, const double* rows,
                                              int stop_on_nonfinite, double* recv, double* t_end) {
    if (!ws || nsteps < 0 || nrecv < 0 || (nrecv > 0 && (!recv_j || !recv_k || !rows || !recv))) return WEDG_ERR_ARGUMENT;
    wedg_problem* p = ws->p;
    const size_t need = (size_t)nsteps * nrecv;
    wedg_status s;
    if ((s = recv_capacity(ws, need))) return s;
    if ((s = recv_rows(ws, nrecv, recv_j, recv_k, rows))) return s;
    if ((s = reset_err(ws))) return s;
    double t = t0;
    RecvArgs ra = recv_args(ws, nrecv);
    for (long it = 0; it < nsteps; ++it) {
        if ((s = rk4_device(ws, t, dt, it + 1, it == 0, it == nsteps - 1))) return s;
        t += dt;
        if (nrecv > 0) {
            ra.step = it;
            HIPOK(launch_diag(p, nullptr, &ra, nullptr, ws->stream));
        }
    }
    if (t_end) *t_end = t;
    if ((s = collect(ws, stop_on_nonfinite))) return s;
    return recv_out(ws, recv, need);
}

extern "C" wedg_status wedg_set_receivers(wedg_workspace* ws, int nrecv, const int* recv_j, const int* recv_k,
                                          const double* rows) {
    if (!ws) return WEDG_ERR_ARGUMENT;
    wedg_status s = recv_rows(ws, nrecv, recv_j, recv_k, rows);
    ws->nrecv_set = s == WEDG_OK ? nrecv : 0;
    return s;
}

extern "C" wedg_status wedg_diagnostics(wedg_workspace* ws, double* E, double* vmax, double* ndmin, int j0,
                                        const double* eta0, double* slice) {
    if (!ws || !E || !vmax || !ndmin || !slice || !eta0) return WEDG_ERR_ARGUMENT;
    wedg_problem* p = ws->p;
    if (!p->dcls) return WEDG_ERR_ARGUMENT;   // wedg_problem_set_diag first
    const int N = p->N, nq = p->nu, nv = p->nv;
    if (j0 < 0 || j0 >= N) return WEDG_ERR_ARGUMENT;
    const size_t n = 3 * (size_t)p->E + (size_t)N * nq + nv;
    HIPOK(hipStreamSynchronize(ws->stream));
    if (!ws->ddiag) {
        void* d = nullptr;
        HIPOK(hipMalloc(&d, n * sizeof(double)));
        ws->allocs.push_back(d);
        ws->ddiag = (double*)d;
    }
    if (!ws->hdiag) HIPOK(hipHostMalloc((void**)&ws->hdiag, n * sizeof(double), hipHostMallocDefault));
    double* dslice = ws->ddiag + 3 * (size_t)p->E;
    double* deta = dslice + (size_t)N * nq;
    std::memcpy(ws->hdiag, eta0, nv * sizeof(double));
    HIPOK(hipMemcpyAsync(deta, ws->hdiag, nv * sizeof(double), hipMemcpyHostToDevice, ws->stream));
    DiagArgs d;
    d.u = ws->yu; d.v = ws->yv; d.N = N; d.E = p->E; d.j0 = j0; d.pad = 0; d.h = 1.0 / N;
    d.tab = p->tab; d.W2d = p->dW2d; d.cls = p->dcls; d.c2 = p->dc2; d.bq = p->dbq; d.eta0 = deta;
    d.E_out = ws->ddiag; d.vmax_out = ws->ddiag + p->E; d.nd_out = ws->ddiag + 2 * (size_t)p->E; d.slice = dslice;
    d.grid = nullptr;
    HIPOK(launch_diag(p, &d, nullptr, nullptr, ws->stream));
    HIPOK(hipMemcpyAsync(ws->hdiag, ws->ddiag, (3 * (size_t)p->E + (size_t)N * nq) * sizeof(double),
                         hipMemcpyDeviceToHost, ws->stream));
    HIPOK(hipStreamSynchronize(ws->stream));
    std::memcpy(E, ws->hdiag, p->E * sizeof(double));
    std::memcpy(vmax, ws->hdiag + p->E, p->E * sizeof(double));
    std::memcpy(ndmin, ws->hdiag + 2 * (size_t)p->E, p->E * sizeof(double));
    std::memcpy(slice, ws->hdiag + 3 * (size_t)p->E, (size_t)N * nq * sizeof(double));
    return WEDG_OK;
}

extern "C" wedg_status wedg_snapshot(wedg_workspace* ws, double* grid) {
    if (!ws || !grid) return WEDG_ERR_ARGUMENT;
    wedg_problem* p = ws->p;
    if (!p->dcls) return WEDG_ERR_ARGUMENT;   // wedg_problem_set_diag first
    const size_t n = (size_t)p->E * p->nu * p->nu;
    HIPOK(hipStreamSynchronize(ws->stream));
    // the grid (E nu^2 values) is formed in the RK4 accumulator, which holds
    // nothing between calls, and read back through the pinned state staging
    double* dg = ws->au;
    DiagArgs d{};
    d.u = ws->yu; d.v = ws->yv; d.N = p->N; d.E = p->E; d.j0 = 0; d.h = 1.0 / p->N;
    d.tab = p->tab; d.W2d = p->dW2d; d.cls = p->dcls; d.c2 = p->dc2; d.bq = p->dbq; d.eta0 = nullptr;
    d.grid = dg;
    HIPOK(launch_diag(p, &d, nullptr, nullptr, ws->stream));
    HIPOK(hipMemcpyAsync(ws->hu, dg, n * sizeof(double), hipMemcpyDeviceToHost, ws->stream));
    HIPOK(hipStreamSynchronize(ws->stream));
    std::memcpy(grid, ws->hu, n * sizeof(double));
    return WEDG_OK;
}

// ---------------------------------------------------------------------------
// Error norms (src/solver.jl:774-836): per-element quadrature sums on the device.

static wedg_status norms_buffers(wedg_workspace* ws) {
    wedg_problem* p = ws->p;
    const size_t n = 4 * (size_t)p->E + 4 * (size_t)p->N * p->nu;
    if (!ws->dnorm) {
        void* d = nullptr;
        HIPOK(hipMalloc(&d, n * sizeof(double)));
        ws->allocs.push_back(d);
        ws->dnorm = (double*)d;
    }
    if (!ws->hnorm) HIPOK(hipHostMalloc((void**)&ws->hnorm, n * sizeof(double), hipHostMallocDefault));
    return WEDG_OK;
}

static NormArgs norm_args(wedg_workspace* ws) {
    wedg_problem* p = ws->p;
    NormArgs a{};
    a.N = p->N; a.E = p->E; a.h = 1.0 / p->N;
    a.tab = p->tab; a.W2d = p->dW2d; a.cls = p->dcls; a.c2 = p->dc2;
    a.out = ws->dnorm;
    return a;
}

static wedg_status norms_out(wedg_workspace* ws, double* elem) {
    const size_t n = 4 * (size_t)ws->p->E;
    HIPOK(hipMemcpyAsync(ws->hnorm, ws->dnorm, n * sizeof(double), hipMemcpyDeviceToHost, ws->stream));
    HIPOK(hipStreamSynchronize(ws->stream));
    std::memcpy(elem, ws->hnorm, n * sizeof(double));
    return WEDG_OK;
}

extern "C" wedg_status wedg_mms_errors(wedg_workspace* ws, double cu, double cv, double tpx, double tpy,
                                       const double* cx, const double* sx, const double* cy, const double* sy,
                                       double* elem) {
    if (!ws || !cx || !sx || !cy || !sy || !elem) return WEDG_ERR_ARGUMENT;
    wedg_problem* p = ws->p;
    if (!p->dcls) return WEDG_ERR_ARGUMENT;   // wedg_problem_set_diag first
    wedg_status s;
    HIPOK(hipStreamSynchronize(ws->stream));
    if ((s = norms_buffers(ws))) return s;
    const size_t m = (size_t)p->N * p->nu;
    double* dt = ws->dnorm + 4 * (size_t)p->E;
    std::memcpy(ws->hnorm, cx, m * sizeof(double));
    std::memcpy(ws->hnorm + m, sx, m * sizeof(double));
    std::memcpy(ws->hnorm + 2 * m, cy, m * sizeof(double));
    std::memcpy(ws->hnorm + 3 * m, sy, m * sizeof(double));
    HIPOK(hipMemcpyAsync(dt, ws->hnorm, 4 * m * sizeof(double), hipMemcpyHostToDevice, ws->stream));
    NormArgs a = norm_args(ws);
    a.u = ws->yu; a.v = ws->yv; a.ub = nullptr; a.vb = nullptr;
    a.cu = cu; a.cv = cv; a.tpx = tpx; a.tpy = tpy;
    a.cx = dt; a.sx = dt + m; a.cy = dt + 2 * m; a.sy = dt + 3 * m;
    HIPOK(launch_diag(p, nullptr, nullptr, &a, ws->stream));
    return norms_out(ws, elem);
}

extern "C" wedg_status wedg_difference_norms(wedg_workspace* ws, const double* uA, const double* vA,
                                             const double* uB, const double* vB, double* elem) {
    if (!ws || !uA || !vA || !uB || !vB || !elem) return WEDG_ERR_ARGUMENT;
    wedg_problem* p = ws->p;
    if (!p->dcls) return WEDG_ERR_ARGUMENT;
    wedg_status s;
    if ((s = norms_buffers(ws))) return s;
    // the two states go to the stage buffers, which hold nothing between calls
    if ((s = stage_in(ws, ws->su, ws->hu, uA, ws->nU()))) return s;
    if ((s = stage_in(ws, ws->sv, ws->hv, vA, ws->nV()))) return s;
    if ((s = stage_in(ws, ws->ou, ws->hu, uB, ws->nU()))) return s;
    if ((s = stage_in(ws, ws->ov, ws->hv, vB, ws->nV()))) return s;
    NormArgs a = norm_args(ws);
    a.u = ws->su; a.v = ws->sv; a.ub = ws->ou; a.vb = ws->ov;
    HIPOK(launch_diag(p, nullptr, nullptr, &a, ws->stream));
    return norms_out(ws, elem);
}

// ---------------------------------------------------------------------------
// Batched instances: the workspaces of one (n_u, n_v, mode) group advance in
// one launch per RK stage; every instance keeps its own tables, state, time,
// step count and failure record, and gets exactly the stage arguments (and so
// the results) its own wedg_advance would give it.

static hipError_t launch_batch(const BatchArgs& b, int grid, int qu, int nv, int which, hipStream_t s, int mode,
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

struct BatchGroup {
    int qu = 0, nv = 0, mode = 0;
    bool two = false;             // the two-lane kernels (every instance's lanes per element is 2)
    int lane = 0;                 // the stream that runs the group
    std::vector<int> inst;        // the group's workspaces (indexes into wedg_batch::ws)
    std::vector<int> order;       // inst by decreasing step count: the active ones are a prefix
    std::vector<int> bfirst;      // first block of order[i]; bfirst[n] = all blocks
    int* dmap = nullptr;          // device: binst [blocks], then bfirst [instances + 1]
};

// One stream of a batch and the stage arguments of its groups for a chunk of
// steps, staged through two pinned buffers.
struct BatchLane {
    hipStream_t stream = nullptr;
    size_t cap = 0;               // KArgs per buffer
    KArgs* hargs[2] = {nullptr, nullptr};
    KArgs* dargs[2] = {nullptr, nullptr};
    hipEvent_t copied[2] = {nullptr, nullptr};
    hipEvent_t done = nullptr;
};

struct wedg_batch {
    std::vector<wedg_workspace*> ws;
    std::vector<BatchGroup> groups;
    std::vector<BatchLane> lanes;
    hipStream_t stream = nullptr; // block maps, timing events, the join of the lanes
    std::vector<void*> allocs;
    long chunk = 1;               // steps per chunk of stage arguments
    int schedule = 0;             // 0: every instance starts at the first step; 1: all end at the last
    int* hmap = nullptr;          // pinned staging of the block maps
    size_t map_ints = 0;
    double last_ms = 0.0;         // device time of the last advance (stream events)
    hipEvent_t t0 = nullptr, t1 = nullptr;
};

constexpr int kMaxBatchLanes = 16;

// Streams of a batch: the hardware queues of a process (GPU_MAX_HW_QUEUES, 4
// by default), so that the lanes run concurrently.
static int batch_lanes(int ngroups) {
    int q = 4;
    if (const char* e = std::getenv("GPU_MAX_HW_QUEUES")) q = std::max(1, std::atoi(e));
    return std::max(1, std::min(ngroups, q));
}

// The FP64 operations of one element and stage without Chebyshev steps (traces,
// volume terms, u-solve, lifts, Chebyshev set-up): the weight of a group when
// the groups are dealt to the lanes.
static double stage_cost(int nu, int nv, int nh) {
    const double q = nu, v = nv, h = nh;
    return 2 * (4 * (q * q + q * q) + 8 * (v * v + q * v)) +
           2 * (2 * (q * q * q + q * q * q) + 2 * (q * v * v + q * q * v) + 2 * (v * q * q + v * v * q) +
                2 * (q * q * q + q * q * q)) +
           2 * q * q * q * q + 4 * (3 * q * (2 * v + q) + 3 * (2 * v * v + q * q)) + 2 * (h * v * v + h * h * v);
}

static void lanes_free(wedg_batch* b) {
    for (auto& l : b->lanes) {
        if (l.stream) (void)hipStreamSynchronize(l.stream);
        for (int i = 0; i < 2; ++i) {
            if (l.hargs[i]) (void)hipHostFree(l.hargs[i]);
            if (l.dargs[i]) (void)hipFree(l.dargs[i]);
            if (l.copied[i]) (void)hipEventDestroy(l.copied[i]);
        }
        if (l.done) (void)hipEventDestroy(l.done);
        if (l.stream) (void)hipStreamDestroy(l.stream);
    }
    b->lanes.clear();
}

// nl streams with stage-argument buffers for chunks of `chunk` steps (0: the
// chunk whose arguments fill about 4 MB over all lanes)
static hipError_t lanes_alloc(wedg_batch* b, int nl, long chunk) {
    const size_t n = b->ws.size();
    b->chunk = chunk > 0 ? chunk : (long)std::max<size_t>(1, ((size_t)4 << 20) / (4 * n * sizeof(KArgs)));
    b->lanes.resize(nl);
    hipError_t he = hipSuccess;
    for (auto& l : b->lanes) {
        l.cap = (size_t)b->chunk * 4 * n;   // any share of the groups
        if (he == hipSuccess) he = hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking);
        for (int i = 0; i < 2 && he == hipSuccess; ++i) {
            he = hipHostMalloc((void**)&l.hargs[i], l.cap * sizeof(KArgs), hipHostMallocDefault);
            void* d = nullptr;
            if (he == hipSuccess) he = hipMalloc(&d, l.cap * sizeof(KArgs));
            if (he == hipSuccess) l.dargs[i] = (KArgs*)d;
            if (he == hipSuccess) he = hipEventCreateWithFlags(&l.copied[i], hipEventDisableTiming);
        }
        if (he == hipSuccess) he = hipEventCreateWithFlags(&l.done, hipEventDisableTiming);
    }
    return he;
}

extern "C" wedg_status wedg_batch_destroy(wedg_batch* b) {
    if (!b) return WEDG_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    lanes_free(b);
    if (b->hmap) (void)hipHostFree(b->hmap);
    if (b->t0) (void)hipEventDestroy(b->t0);
    if (b->t1) (void)hipEventDestroy(b->t1);
    for (void* a : b->allocs) (void)hipFree(a);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return WEDG_OK;
}

extern "C" wedg_status wedg_batch_create(wedg_workspace* const* ws, int n, wedg_batch** out) {
    if (!ws || !out || n <= 0) return WEDG_ERR_ARGUMENT;
    for (int i = 0; i < n; ++i) {
        if (!ws[i]) return WEDG_ERR_ARGUMENT;
        for (int j = 0; j < i; ++j)
            if (ws[j] == ws[i]) return WEDG_ERR_ARGUMENT;   // one state per instance
    }
    auto* b = new wedg_batch();
    b->ws.assign(ws, ws + n);
    for (int i = 0; i < n; ++i) {
        // a group shares the kernels of its instances' own calls: the same
        // (n_u, n_v, mode) and lanes per element
        const wedg_problem* p = ws[i]->p;
        const bool two = ws[i]->two();
        auto it = std::find_if(b->groups.begin(), b->groups.end(), [&](const BatchGroup& g) {
            return g.qu == p->qu && g.nv == p->nv && g.mode == p->mode && g.two == two;
        });
        if (it == b->groups.end()) {
            b->groups.emplace_back();
            it = b->groups.end() - 1;
            it->qu = p->qu; it->nv = p->nv; it->mode = p->mode; it->two = two;
        }
        it->inst.push_back(i);
    }
    hipError_t he = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (he == hipSuccess) he = lanes_alloc(b, batch_lanes((int)b->groups.size()), 0);
    for (auto& g : b->groups) {
        size_t blocks = 0;
        for (int i : g.inst) blocks += ws[i]->nwaves();
        const size_t ints = blocks + g.inst.size() + 1;
        b->map_ints += ints;
        void* d = nullptr;
        if (he == hipSuccess) he = hipMalloc(&d, ints * sizeof(int));
        if (he == hipSuccess) { b->allocs.push_back(d); g.dmap = (int*)d; }
    }
    if (he == hipSuccess) he = hipHostMalloc((void**)&b->hmap, b->map_ints * sizeof(int), hipHostMallocDefault);
    if (he == hipSuccess) he = hipEventCreate(&b->t0);
    if (he == hipSuccess) he = hipEventCreate(&b->t1);
    if (he != hipSuccess) { wedg_batch_destroy(b); return WEDG_ERR_HIP; }
    *out = b;
    return WEDG_OK;
}

// The batched run; recv (may be null): per instance, null or room for
// nsteps[i] x (its wedg_set_receivers count) values of v_h at its receivers
// after every step, [step][receiver].
static wedg_status batch_run(wedg_batch* b, const double* t0, const double* dt, const long* nsteps,
                             int stop_on_nonfinite, double* t_end, int* status, double* const* recv) {
    if (!b || !t0 || !dt || !nsteps) return WEDG_ERR_ARGUMENT;
    const int n = (int)b->ws.size();
    for (int i = 0; i < n; ++i)
        if (nsteps[i] < 0) return WEDG_ERR_ARGUMENT;
    auto records = [&](int i) { return recv && recv[i] && b->ws[i]->nrecv_set > 0; };
    for (int i = 0; i < n; ++i)
        if (records(i)) {
            const wedg_status sc = recv_capacity(b->ws[i], (size_t)nsteps[i] * b->ws[i]->nrecv_set);
            if (sc) return sc;
        }
    for (const auto& l : b->lanes)
        if (!l.stream || !l.dargs[0] || !l.dargs[1] || !l.hargs[0] || !l.hargs[1]) return WEDG_ERR_HIP;
    // the block maps and kernels follow the lanes per element of batch_create
    for (const auto& g : b->groups)
        for (int i : g.inst)
            if (b->ws[i]->two() != g.two) return WEDG_ERR_ARGUMENT;
    wedg_status s;
    for (int i = 0; i < n; ++i) {
        if ((s = reset_err(b->ws[i]))) return s;
        HIPOK(hipStreamSynchronize(b->ws[i]->stream));
    }
    // block maps: the instances of a group by decreasing step count
    HIPOK(hipStreamSynchronize(b->stream));   // the pinned maps may be the source of a copy in flight
    size_t mo = 0;
    long maxsteps = 0;
    for (auto& g : b->groups) {
        g.order = g.inst;
        std::stable_sort(g.order.begin(), g.order.end(), [&](int x, int y) { return nsteps[x] > nsteps[y]; });
        g.bfirst.assign(1, 0);
        int* m = b->hmap + mo;
        int blk = 0;
        for (size_t r = 0; r < g.order.size(); ++r) {
            const int nw = (int)b->ws[g.order[r]]->nwaves();
            for (int w = 0; w < nw; ++w) m[blk + w] = (int)r;
            blk += nw;
            g.bfirst.push_back(blk);
        }
        std::copy(g.bfirst.begin(), g.bfirst.end(), m + blk);
        const size_t ints = blk + g.bfirst.size();
        HIPOK(hipMemcpyAsync(g.dmap, m, ints * sizeof(int), hipMemcpyHostToDevice, b->stream));
        mo += ints;
        if (!g.order.empty()) maxsteps = std::max(maxsteps, nsteps[g.order[0]]);
    }
    // groups to lanes, largest first, each to the least loaded lane; the weight
    // of a group is its cost per element and stage (the stage time of a wave,
    // one element per lane, half that with two) times its largest step count
    {
        std::vector<std::pair<double, int>> w;
        for (size_t g = 0; g < b->groups.size(); ++g) {
            const wedg_problem* p0 = b->ws[b->groups[g].order[0]]->p;
            const double pair = b->groups[g].two ? 0.5 : 1.0;   // two lanes share an element's work
            w.push_back({stage_cost(p0->nu, p0->nv, p0->nh) * pair * (double)nsteps[b->groups[g].order[0]], (int)g});
        }
        std::stable_sort(w.begin(), w.end(), [](auto& x, auto& y) { return x.first > y.first; });
        std::vector<double> load(b->lanes.size(), 0.0);
        for (auto& x : w) {
            const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            load[l] += x.first;
            b->groups[x.second].lane = l;
        }
    }
    HIPOK(hipEventRecord(b->t0, b->stream));
    for (auto& l : b->lanes) HIPOK(hipStreamWaitEvent(l.stream, b->t0, 0));
    // instance i runs its steps at the batch's steps o[i] .. o[i] + nsteps[i] - 1;
    // in both schedules the active instances of a group are a prefix of its order
    std::vector<long> o(n, 0);
    if (b->schedule == 1)
        for (int i = 0; i < n; ++i) o[i] = maxsteps - nsteps[i];
    auto active = [&](int i, long it) { return it >= o[i] && it < o[i] + nsteps[i]; };
    std::vector<double> t(t0, t0 + n);   // every instance's time: t += dt after each of its steps
    std::vector<double> tc;              // [instance][step of the chunk]: the time at the step's start
    struct Launch { int g, which, grid; size_t off; int inst; long step; };
    std::vector<Launch> L;
    int buf = 0;
    for (long c0 = 0; c0 < maxsteps; c0 += b->chunk, buf ^= 1) {
        const long c1 = std::min(maxsteps, c0 + b->chunk), len = c1 - c0;
        tc.resize((size_t)n * len);
        for (int i = 0; i < n; ++i)
            for (long j = 0; j < len; ++j) {
                tc[(size_t)i * len + j] = t[i];
                if (active(i, c0 + j)) t[i] += dt[i];
            }
        for (size_t li = 0; li < b->lanes.size(); ++li) {
            BatchLane& l = b->lanes[li];
            HIPOK(hipEventSynchronize(l.copied[buf]));   // the pinned buffer's last copy has run
            KArgs* H = l.hargs[buf];
            size_t k = 0;
            L.clear();
            for (long it = c0; it < c1; ++it)
                for (int st = 1; st <= 4; ++st)
                    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
                        const BatchGroup& g = b->groups[gi];
                        if (g.lane != (int)li) continue;
                        size_t na = 0;
                        while (na < g.order.size() && active(g.order[na], it)) ++na;
                        if (na == 0) continue;
                        const size_t off = k;
                        bool starts = false;   // an instance takes its first step here
                        for (size_t r = 0; r < na; ++r) {
                            const int i = g.order[r];
                            const long s = it - o[i];   // the instance's own step
                            starts = starts || s == 0;
                            H[k++] = rk4_args(b->ws[i], st, tc[(size_t)i * len + (it - c0)], dt[i], s + 1,
                                              s == nsteps[i] - 1);
                        }
                        // the first stage of an instance's run needs k_prep; for the instances
                        // already running it rewrites the values their last stage formed
                        if (st == 1 && starts) L.push_back({(int)gi, 0, g.bfirst[na], off, -1, 0});
                        L.push_back({(int)gi, 1, g.bfirst[na], off, -1, 0});
                        if (st == 4)   // v_h at the receivers after the step
                            for (size_t r = 0; r < na; ++r)
                                if (records(g.order[r]))
                                    L.push_back({(int)gi, 2, 0, 0, g.order[r], it - o[g.order[r]]});
                    }
            if (k == 0) continue;
            HIPOK(hipMemcpyAsync(l.dargs[buf], H, k * sizeof(KArgs), hipMemcpyHostToDevice, l.stream));
            HIPOK(hipEventRecord(l.copied[buf], l.stream));
            for (const Launch& x : L) {
                const BatchGroup& g = b->groups[x.g];
                if (x.which == 2) {
                    wedg_workspace* w = b->ws[x.inst];
                    RecvArgs ra = recv_args(w, w->nrecv_set);
                    ra.step = x.step;
                    HIPOK(launch_diag(w->p, nullptr, &ra, nullptr, l.stream));
                    continue;
                }
                BatchArgs ba;
                ba.args = l.dargs[buf] + x.off;
                ba.binst = g.dmap;
                ba.bfirst = g.dmap + g.bfirst.back();
                HIPOK(launch_batch(ba, x.grid, g.qu, g.nv, x.which, l.stream, g.mode, g.two));
            }
        }
    }
    for (auto& l : b->lanes) {
        HIPOK(hipEventRecord(l.done, l.stream));
        HIPOK(hipStreamWaitEvent(b->stream, l.done, 0));
    }
    HIPOK(hipEventRecord(b->t1, b->stream));
    HIPOK(hipStreamSynchronize(b->stream));
    float ms = 0.f;
    b->last_ms = hipEventElapsedTime(&ms, b->t0, b->t1) == hipSuccess ? ms : 0.0;
    wedg_status first = WEDG_OK;
    for (int i = 0; i < n; ++i) {
        wedg_status si = collect(b->ws[i], stop_on_nonfinite);
        if (si == WEDG_OK && records(i)) si = recv_out(b->ws[i], recv[i], (size_t)nsteps[i] * b->ws[i]->nrecv_set);
        if (status) status[i] = si;
        if (t_end) t_end[i] = t[i];
        if (si != WEDG_OK && first == WEDG_OK) first = si;
    }
    return first;
}

extern "C" wedg_status wedg_batch_advance(wedg_batch* b, const double* t0, const double* dt, const long* nsteps,
                                          int stop_on_nonfinite, double* t_end, int* status) {
    return batch_run(b, t0, dt, nsteps, stop_on_nonfinite, t_end, status, nullptr);
}

extern "C" wedg_status wedg_batch_advance_receivers(wedg_batch* b, const double* t0, const double* dt,
                                                    const long* nsteps, int stop_on_nonfinite, double* t_end,
                                                    int* status, double* const* recv) {
    if (!recv) return WEDG_ERR_ARGUMENT;
    return batch_run(b, t0, dt, nsteps, stop_on_nonfinite, t_end, status, recv);
}

extern "C" wedg_status wedg_batch_configure(wedg_batch* b, int lanes, long chunk_steps, int schedule) {
    if (!b || lanes < 1 || lanes > kMaxBatchLanes || chunk_steps < 0 || schedule < 0 || schedule > 1)
        return WEDG_ERR_ARGUMENT;
    b->schedule = schedule;
    HIPOK(hipStreamSynchronize(b->stream));
    lanes_free(b);
    if (lanes_alloc(b, lanes, chunk_steps) != hipSuccess) {
        lanes_free(b);
        (void)lanes_alloc(b, 1, 0);   // leave a usable batch if the larger set did not fit
        return WEDG_ERR_HIP;
    }
    return WEDG_OK;
}

extern "C" wedg_status wedg_batch_config(const wedg_batch* b, int* lanes, long* chunk_steps, int* schedule) {
    if (!b || !lanes || !chunk_steps || !schedule) return WEDG_ERR_ARGUMENT;
    *lanes = (int)b->lanes.size();
    *chunk_steps = b->chunk;
    *schedule = b->schedule;
    return WEDG_OK;
}

extern "C" wedg_status wedg_batch_time(const wedg_batch* b, double* ms) {
    if (!b || !ms) return WEDG_ERR_ARGUMENT;
    *ms = b->last_ms;
    return WEDG_OK;
}
