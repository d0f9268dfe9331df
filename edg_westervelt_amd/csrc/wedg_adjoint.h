// Tangent-linear and adjoint (vector-Jacobian product) kernels of one RK4 stage
// of the source-mode right-hand side with per-element coefficients c², α, β
// (gfx950, FP64), and the boundary misfit of u_h.
//
// The forward map is the reference's edg_rhs! in mode = :source
// (src/rhs.jl:490-550): zero boundary data and no u₀ terms, faces with one-sided
// coefficients (src/rhs.jl:663-700), boundary kinds of src/rhs.jl:74-96, lifts
// and volume terms (src/rhs.jl:317-350, 421-473), the u-equation with the pinned
// c²-stiffness (src/rhs.jl:361-382) and the mass system of the v-equation
// (src/rhs.jl:778-790, 886-939). The reference has no adjoint: these kernels
// differentiate its discrete map. The mass solve X = M⁻¹R is differentiated as
// the exact solve, dX = M⁻¹(dR − dM X), not through the iteration; its adjoint
// λ = M⁻¹X̄ is a second solve with the same Chebyshev iteration (M is
// symmetric), to the same 1e-14 residual bound, each element with its own
// a-priori step count.
//
// One lane owns one element, as in k_stage. The VJP of a stage runs in two
// kernels: k_adj_a works on the element alone (the mass and u-solve adjoints,
// the volume terms, the lifts) and writes the adjoints of the element's face
// slots; k_adj_b gathers, face by face, the slot adjoints of the element and of
// its neighbour, and turns them into the adjoints of the element's traces and
// coefficients (each element takes its own share of a face, so no slot and no
// gradient entry has two writers), then does the RK4 bookkeeping of the stage.
#pragma once

#include "wedg_kernels.h"

namespace wedg {

// Arguments of the adjoint, tangent-linear and misfit kernels.
struct AdjArgs {
    int N, E, stage, pad0;
    double two_h, four_hh, theta, tau, lu, lv, cheb_L, hw;   // hw = h / 2 (face weight scale)
    int kind[4];
    double gam[4], eta[4], sig[4], kb[4], kap[4];
    const double* tab;      // the forward table (Dims offsets)
    const double* atab;     // W2d (NQ x NQ, unit coefficient), K = S^T diag(w) S (NU x NU), lambda (NU)
    const double* c2; const double* al; const double* be;    // per element [E]
    const double* yu; const double* yv;                      // the stage state Y_s [m][E]
    const double* ku; const double* kv;                      // F(Y_s)
    const double* kbu; const double* kbv;                    // k_adj_a: cotangent of F(Y_s), times kscale
    double kscale, ca, cb, pad1;                             // k_adj_b: next cotangent = ca ybar + cb Ybar_s
    double* slot;           // slot adjoints [side][A, D, F][i][E], written by k_adj_a
    const double* slotr;    // the slot adjoints k_adj_b reads (the stage's k_adj_a wrote them)
    double* pu; double* pv; // k_adj_a's share of the state cotangent [m][E]
    double* gc2; double* gbe;                                // dJ/dc2, dJ/dbeta [E], accumulated
    double* ybu; double* ybv;                                // ybar of the step; stage 1 replaces it
    double* accu; double* accv;                              // ybar + sum of the stages' Ybar
    double* nkbu; double* nkbv;                              // the next cotangent
    // tangent-linear inputs and outputs
    const double* dc2; const double* dbe; const double* du; const double* dv;
    double* outu; double* outv;
    // misfit
    const double* tgt;      // target u_h at the bottom face points [k][i] (or null)
    double* tr;             // traces out [k][i] (or null)
    double* jp;             // 1/2 w r^2 per point [k][i] (or null)
    int* fail;              // set to 1 when a mass matrix is not positive definite
};

// The kernels' AdjArgs read through the kernarg segment pointer in the constant
// address space, opaque to the optimizer (as kargs() for the forward kernels):
// each phase fetches it anew, so fields are loaded (scalar) where they are used
// instead of being held in SGPRs across the whole kernel.
// A kernel that runs two stages' parts (k_adj_ba) takes two AdjArgs; W selects one.
// W = 2, 3: a batched launch (a BatchArgs whose args point to AdjArgs pairs,
// [instance][2]): the block's instance selects its pair, W - 2 the member.
#if defined(__HIP_DEVICE_COMPILE__)
typedef const WEDG_CONST_AS AdjArgs CAdjArgs;
template <int W = 0>
__device__ __forceinline__ CAdjArgs& adjargs() {
    CAdjArgs* p;
    if constexpr (W >= 2) p = (CAdjArgs*)(const void*)bargs()->args + 2 * block_instance() + (W - 2);
    else p = (CAdjArgs*)__builtin_amdgcn_kernarg_segment_ptr() + W;
    asm volatile("" : "+s"(p));
    return *p;
}
#else
typedef const AdjArgs CAdjArgs;
extern const AdjArgs* wedg_emu_adjargs;   // set by a host build that runs the kernels on the CPU
template <int W = 0>
__device__ __forceinline__ CAdjArgs& adjargs() { return wedg_emu_adjargs[W]; }
#endif

// ---------------------------------------------------------------------------
// Small helpers.

template <int n>
__device__ __forceinline__ void ld(const double* p, int E, int e, double (&x)[n]) {
#pragma unroll
    for (int m = 0; m < n; ++m) x[m] = p[(long)m * E + e];
}

template <int n>
__device__ __forceinline__ void st(double* p, int E, int e, const double (&x)[n]) {
#pragma unroll
    for (int m = 0; m < n; ++m) p[(long)m * E + e] = x[m];
}

// Adjoint of side_trace: X[a, b] += E[a] sum_i P[i, b] o[i] (B, T), or the
// transposed index pair (L, R); E is the row of the side's end.
template <class D, int n, bool DER, int sd>
__device__ __forceinline__ void side_trace_T(cdbl* T, int oEr, const double (&o)[D::NQ], double (&X)[n * n]) {
    constexpr int NQ = D::NQ;
    constexpr bool second = sd >= 2, plus_end = (sd & 1) != 0;
    double z[n];
    cols_sym<NQ, n, false>(T + D::oP, o, z);
#pragma unroll
    for (int a = 0; a < n; ++a) {
        const double r = T[oEr + a];
        const double Ea = plus_end ? r : ((a & 1) ? (DER ? r : -r) : (DER ? -r : r));
#pragma unroll
        for (int b = 0; b < n; ++b) {
            if (second) X[b + n * a] = fma(Ea, z[b], X[b + n * a]);
            else X[a + n * b] = fma(Ea, z[b], X[a + n * b]);
        }
    }
}

// out = A^-1 b with A = c2 (I (x) K + K (x) I) + e0 e0^T, K = Q diag(lam) Q^T
// (the pinned stiffness of src/solver.jl:528-531 for c^2 constant on the element).
template <class D>
__device__ __forceinline__ void usolve(cdbl* Q, cdbl* lam, double c2, const double (&b)[D::NU * D::NU],
                                       double (&out)[D::NU * D::NU]) {
    constexpr int NU = D::NU;
    double Y1[NU * NU], Y2[NU * NU];
#pragma unroll
    for (int bb = 0; bb < NU; ++bb)
#pragma unroll
        for (int i = 0; i < NU; ++i) {
            double s = 0.0;
#pragma unroll
            for (int a = (i & 1); a < NU; a += 2) s = fma(Q[a + NU * i], b[a + NU * bb], s);
            Y1[i + NU * bb] = s;
        }
#pragma unroll
    for (int i = 0; i < NU; ++i)
#pragma unroll
        for (int j = 0; j < NU; ++j) {
            double s = 0.0;
#pragma unroll
            for (int bb = (j & 1); bb < NU; bb += 2) s = fma(Y1[i + NU * bb], Q[bb + NU * j], s);
            const double den = c2 * (lam[i] + lam[j]) + ((i == 0 && j == 0) ? 1.0 : 0.0);
            Y2[i + NU * j] = s / den;
        }
#pragma unroll
    for (int j = 0; j < NU; ++j)
#pragma unroll
        for (int a = 0; a < NU; ++a) {
            double s = 0.0;
#pragma unroll
            for (int i = (a & 1); i < NU; i += 2) s = fma(Q[a + NU * i], Y2[i + NU * j], s);
            Y1[a + NU * j] = s;
        }
#pragma unroll
    for (int bb = 0; bb < NU; ++bb)
#pragma unroll
        for (int a = 0; a < NU; ++a) {
            double s = 0.0;
#pragma unroll
            for (int j = (bb & 1); j < NU; j += 2) s = fma(Y1[a + NU * j], Q[bb + NU * j], s);
            out[a + NU * bb] = s;
        }
}

// <b, K X + X K> for the 1D stiffness K (symmetric): b^T (I (x) K + K (x) I) X
template <class D>
__device__ __forceinline__ double stiff_dot(cdbl* K, const double (&b)[D::NU * D::NU],
                                            const double (&X)[D::NU * D::NU]) {
    constexpr int NU = D::NU;
    double s = 0.0;
#pragma unroll
    for (int bb = 0; bb < NU; ++bb)
#pragma unroll
        for (int a = 0; a < NU; ++a) {
            double t = 0.0;
#pragma unroll
            for (int c = 0; c < NU; ++c) {
                t = fma(K[a + NU * c], X[c + NU * bb], t);
                t = fma(X[a + NU * c], K[c + NU * bb], t);
            }
            s = fma(b[a + NU * bb], t, s);
        }
    return s;
}

// ---------------------------------------------------------------------------
// The mass operator of an element with its weights G (NH x NH) in LDS,
// Gl[(r NH + q) kWave + lane].

template <class D>
__device__ __forceinline__ void kapply(cdbl* Ph, const double* Gl, int lane, const double (&d)[D::NV * D::NV],
                                       double (&out)[D::NV * D::NV]) {
    constexpr int NH = D::NH, NP = D::NP;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        double Gp[NH], Gq[NH];
#pragma unroll
        for (int q = 0; q < NH; ++q) {
            Gp[q] = Gl[(p * NH + q) * kWave + lane];
            Gq[q] = Gl[((NH - 1 - p) * NH + q) * kWave + lane];
        }
        kd_pair<D>(Ph, p, d, Gp, Gq, out);
    }
    if (NH & 1) {
        double Gm[NH];
#pragma unroll
        for (int q = 0; q < NH; ++q) Gm[q] = Gl[((NH / 2) * NH + q) * kWave + lane];
        kd_mid<D>(Ph, d, Gm, out);
    }
}

// v at the high-order points (row r = p, NH-1-p) of the modal V, with f(r, q, value)
template <class D, class F>
__device__ __forceinline__ void for_points(cdbl* Ph, const double (&V)[D::NV * D::NV], F f) {
    constexpr int NH = D::NH;
#pragma unroll
    for (int p = 0; p < (NH + 1) / 2; ++p) {
        const bool mid = (NH & 1) && p == NH / 2;
        double yp[NH], yq[NH];
        vq_rows<D>(Ph, p, mid, V, yp, yq);
#pragma unroll
        for (int q = 0; q < NH; ++q) {
            f(p, q, yp[q]);
            if (!mid) f(NH - 1 - p, q, yq[q]);
        }
    }
}

// G = 2 beta v_q W_h into LDS and the Chebyshev interval of the element
// (src/rhs.jl:889-907); returns its a-priori step count, 0 when 1 - 2 beta v_q <= 0.
template <class D>
__device__ __forceinline__ int mass_setup(cdbl* T, double beta, const double (&V)[D::NV * D::NV], double* Gl,
                                          int lane, double L, double& th, double& de, double& s1) {
    constexpr int NH = D::NH;
    double gmin = 1.0 / 0.0, gmax = -1.0 / 0.0;
    bool nan = false;
    for_points<D>(T + D::oPh, V, [&](int r, int q, double v) {
        const double g = v * (2.0 * beta);
        gmin = fmin(gmin, g);
        gmax = fmax(gmax, g);
        nan = nan || (g != g);
        Gl[(r * NH + q) * kWave + lane] = g * T[D::oWh + r + NH * q];
    });
    return cheb_interval(gmin, gmax, nan, L, th, de, s1);
}

// X = M^-1 R by n Chebyshev steps (Saad, Alg. 12.1; the recurrence of
// src/rhs.jl:920-939), with M = I - K(G).
template <class D>
__device__ __forceinline__ void cheb_solve(cdbl* Ph, const double* Gl, int lane, int n, double th, double de,
                                           double s1, const double (&R)[D::NV * D::NV],
                                           double (&X)[D::NV * D::NV]) {
    constexpr int nV = D::NV * D::NV;
    double r[nV], d[nV];
#pragma unroll
    for (int m = 0; m < nV; ++m) { X[m] = 0.0; r[m] = R[m]; d[m] = r[m] / th; }
    double rho = 1.0 / s1;
    for (int step = 1; step <= n; ++step) {
#pragma unroll
        for (int m = 0; m < nV; ++m) X[m] += d[m];
        if (step == n) break;
#pragma unroll
        for (int m = 0; m < nV; ++m) r[m] = r[m] - d[m];
        kapply<D>(Ph, Gl, lane, d, r);
        const double rhon = 1.0 / (2.0 * s1 - rho);
        const double c1 = rhon * rho, c2 = 2.0 * rhon / de;
#pragma unroll
        for (int m = 0; m < nV; ++m) d[m] = c1 * d[m] + c2 * r[m];
        rho = rhon;
    }
}

// ---------------------------------------------------------------------------
// Faces. Side s of element e = j + N k: B (0), T (1), L (2), R (3); the element
// is the "+" side of the face on B and L. The neighbour across s and the face's
// physical-boundary flag.
__device__ __forceinline__ void side_nb(int s, int e, int j, int k, int N, int& nb, bool& bnd) {
    if (s == 0)      { nb = e - 1; bnd = j == 0; }
    else if (s == 1) { nb = e + 1; bnd = j == N - 1; }
    else if (s == 2) { nb = e - N; bnd = k == 0; }
    else             { nb = e + N; bnd = k == N - 1; }
    if (bnd) nb = e;
}

// Interior face point, one-sided coefficients, K = 0 (src/rhs.jl:683-696): the
// slot values of the element on the side `plus` (true: "+").
template <class KA>
__device__ __forceinline__ void face_slots(const KA& a, bool plus, double vm, double vp, double dum,
                                           double dup, double dvm, double dvp, double c2m, double c2p,
                                           double afm, double afp, double& A, double& Dd, double& Ff) {
    const double th = a.theta, tau = a.tau, two_h = a.two_h;
    const double J = two_h * (c2m * dum - c2p * dup + afm * dvm - afp * dvp);
    const double G = two_h * ((1 - th) * (c2m * dum + afm * dvm) + th * (c2p * dup + afp * dvp));
    const double pen = (0.5 * (c2m + c2p)) * a.lu + (0.5 * (afm + afp)) * a.lv;
    const double s = vm - vp;
    if (plus) {
        const double dvsp = th * s - tau * J;
        A = pen * s - G; Dd = -afp * dvsp; Ff = c2p * dvsp;
    } else {
        const double dvsm = (1 - th) * (vp - vm) - tau * J;
        A = G - pen * s; Dd = afm * dvsm; Ff = -c2m * dvsm;
    }
}

// Tangent of face_slots in the traces and c2 of both sides (alpha fixed).
template <class KA>
__device__ __forceinline__ void face_slots_tl(const KA& a, bool plus, double vm, double vp, double dum,
                                              double dup, double dvm, double dvp, double c2m, double c2p,
                                              double afm, double afp, double tvm, double tvp, double tdum,
                                              double tdup, double tdvm, double tdvp, double tc2m, double tc2p,
                                              double& A, double& Dd, double& Ff) {
    const double th = a.theta, tau = a.tau, two_h = a.two_h;
    const double J = two_h * (c2m * dum - c2p * dup + afm * dvm - afp * dvp);
    const double tJ = two_h * (tc2m * dum + c2m * tdum - tc2p * dup - c2p * tdup + afm * tdvm - afp * tdvp);
    const double tG = two_h * ((1 - th) * (tc2m * dum + c2m * tdum + afm * tdvm) +
                               th * (tc2p * dup + c2p * tdup + afp * tdvp));
    const double pen = (0.5 * (c2m + c2p)) * a.lu + (0.5 * (afm + afp)) * a.lv;
    const double tpen = (0.5 * (tc2m + tc2p)) * a.lu;
    const double s = vm - vp, ts = tvm - tvp;
    if (plus) {
        const double dvsp = th * s - tau * J, tdvsp = th * ts - tau * tJ;
        A = tpen * s + pen * ts - tG; Dd = -afp * tdvsp; Ff = tc2p * dvsp + c2p * tdvsp;
    } else {
        const double dvsm = (1 - th) * (vp - vm) - tau * J, tdvsm = -(1 - th) * ts - tau * tJ;
        A = tG - tpen * s - pen * ts; Dd = afm * tdvsm; Ff = -(tc2m * dvsm + c2m * tdvsm);
    }
}

// Adjoint of face_slots of both sides of a face point: from the slot adjoints of
// the "+" (Abp, Dbp, Fbp) and "-" (Abm, Dbm, Fbm) elements, the adjoints of the
// traces and c2 of the side `plus` (its own share).
template <class KA>
__device__ __forceinline__ void face_slots_T(const KA& a, bool plus, double vm, double vp, double dum,
                                             double dup, double dvm, double dvp, double c2m, double c2p,
                                             double afm, double afp, double Abp, double Dbp, double Fbp,
                                             double Abm, double Dbm, double Fbm, double& vb, double& dub,
                                             double& dvb, double& c2b) {
    const double th = a.theta, tau = a.tau, two_h = a.two_h;
    const double J = two_h * (c2m * dum - c2p * dup + afm * dvm - afp * dvp);
    const double pen = (0.5 * (c2m + c2p)) * a.lu + (0.5 * (afm + afp)) * a.lv;
    const double s = vm - vp;
    const double dvsp = th * s - tau * J;
    const double dvsm = (1 - th) * (vp - vm) - tau * J;
    const double dvspb = -afp * Dbp + c2p * Fbp;
    const double dvsmb = afm * Dbm - c2m * Fbm;
    const double sb = (Abp - Abm) * pen + th * dvspb - (1 - th) * dvsmb;
    const double penb = (Abp - Abm) * s;
    const double Gb = Abm - Abp;
    const double Jb = -tau * (dvspb + dvsmb);
    if (plus) {
        vb = -sb;
        dub = -two_h * c2p * Jb + two_h * th * c2p * Gb;
        dvb = -two_h * afp * Jb + two_h * th * afp * Gb;
        c2b = dvsp * Fbp + 0.5 * a.lu * penb - two_h * dup * Jb + two_h * th * dup * Gb;
    } else {
        vb = sb;
        dub = two_h * c2m * Jb + two_h * (1 - th) * c2m * Gb;
        dvb = two_h * afm * Jb + two_h * (1 - th) * afm * Gb;
        c2b = -dvsm * Fbm + 0.5 * a.lu * penb + two_h * dum * Jb + two_h * (1 - th) * dum * Gb;
    }
}

// Physical boundary point of side s with zero data (src/rhs.jl:74-96, 738-763):
// the slots; returns false when the weight 1 - k beta v_h is not positive.
template <class KA>
__device__ __forceinline__ bool bnd_slots(const KA& a, int s, double vt, double du, double dv, double c2,
                                          double af, double bf, double& A, double& Dd, double& Ff) {
    const double sgn = (s == 0 || s == 2) ? -1.0 : 1.0;
    const double cf = sqrt(c2);
    const double gun = sgn * (a.two_h * du), gvn = sgn * (a.two_h * dv);
    double F1, dvv;
    if (a.kind[s] == 3) {
        const double rho = c2 * gun + af * gvn;
        F1 = 0.0;
        dvv = -(a.kap[s] / cf) * rho;
    } else {
        const double wgt = 1.0 - a.kb[s] * bf * vt;
        if (!(wgt > 0.0)) return false;
        const double sq = sqrt(wgt);
        const double et = a.eta[s], g = a.gam[s], sg = a.sig[s];
        const double rho = et * (c2 * gun + af * gvn) + g * cf * sq * vt;
        F1 = c2 * gun + af * gvn - (et + sg * g) * rho;
        dvv = -(g - sg * et) / (cf * sq) * rho;
    }
    A = F1; Dd = sgn * (af * dvv); Ff = -sgn * (c2 * dvv);
    return true;
}

// Tangent of bnd_slots in (vt, du, dv, c2, bf).
template <class KA>
__device__ __forceinline__ void bnd_slots_tl(const KA& a, int s, double vt, double du, double dv, double c2,
                                             double af, double bf, double tvt, double tdu, double tdv, double tc2,
                                             double tbf, double& A, double& Dd, double& Ff) {
    const double sgn = (s == 0 || s == 2) ? -1.0 : 1.0;
    const double cf = sqrt(c2), tcf = tc2 / (2.0 * cf);
    const double gun = sgn * (a.two_h * du), gvn = sgn * (a.two_h * dv);
    const double tgun = sgn * (a.two_h * tdu), tgvn = sgn * (a.two_h * tdv);
    double tF1, dvv, tdvv;
    if (a.kind[s] == 3) {
        const double rho = c2 * gun + af * gvn;
        const double trho = tc2 * gun + c2 * tgun + af * tgvn;
        const double kp = a.kap[s];
        tF1 = 0.0;
        dvv = -(kp / cf) * rho;
        tdvv = -(kp / cf) * trho + kp * rho * tcf / (cf * cf);
    } else {
        const double kb = a.kb[s], et = a.eta[s], g = a.gam[s], sg = a.sig[s];
        const double wgt = 1.0 - kb * bf * vt, twgt = -kb * (tbf * vt + bf * tvt);
        const double sq = sqrt(wgt), tsq = twgt / (2.0 * sq);
        const double rho = et * (c2 * gun + af * gvn) + g * cf * sq * vt;
        const double trho = et * (tc2 * gun + c2 * tgun + af * tgvn) + g * (tcf * sq * vt + cf * tsq * vt + cf * sq * tvt);
        tF1 = tc2 * gun + c2 * tgun + af * tgvn - (et + sg * g) * trho;
        const double kp = g - sg * et, den = cf * sq, tden = tcf * sq + cf * tsq;
        dvv = -kp / den * rho;
        tdvv = -kp / den * trho + kp * rho * tden / (den * den);
    }
    A = tF1; Dd = sgn * (af * tdvv); Ff = -sgn * (tc2 * dvv + c2 * tdvv);
}

// Adjoint of bnd_slots: from (Ab, Db, Fb), the adjoints of (vt, du, dv, c2, bf).
template <class KA>
__device__ __forceinline__ void bnd_slots_T(const KA& a, int s, double vt, double du, double dv, double c2,
                                            double af, double bf, double Ab, double Db, double Fb, double& vtb,
                                            double& dub, double& dvb, double& c2b, double& bfb) {
    const double sgn = (s == 0 || s == 2) ? -1.0 : 1.0;
    const double cf = sqrt(c2);
    const double gun = sgn * (a.two_h * du), gvn = sgn * (a.two_h * dv);
    const double dvvb = sgn * af * Db - sgn * c2 * Fb;
    double cfb = 0.0, gunb, gvnb;
    vtb = 0.0; bfb = 0.0;
    if (a.kind[s] == 3) {
        const double kp = a.kap[s];
        const double rho = c2 * gun + af * gvn;
        const double dvv = -(kp / cf) * rho;
        c2b = -sgn * dvv * Fb;
        const double rhob = -(kp / cf) * dvvb;
        cfb = (kp / (cf * cf)) * rho * dvvb;
        gunb = c2 * rhob;
        gvnb = af * rhob;
        c2b += gun * rhob;
    } else {
        const double kb = a.kb[s], et = a.eta[s], g = a.gam[s], sg = a.sig[s];
        const double sq = sqrt(1.0 - kb * bf * vt);
        const double rho = et * (c2 * gun + af * gvn) + g * cf * sq * vt;
        const double kp = g - sg * et;
        const double dvv = -kp / (cf * sq) * rho;
        c2b = -sgn * dvv * Fb;
        const double rhob = -(et + sg * g) * Ab - kp / (cf * sq) * dvvb;
        cfb = kp * rho / (cf * cf * sq) * dvvb;
        double sqb = kp * rho / (cf * sq * sq) * dvvb;
        c2b += gun * Ab + et * gun * rhob;
        gunb = c2 * Ab + et * c2 * rhob;
        gvnb = af * Ab + et * af * rhob;
        cfb += g * sq * vt * rhob;
        sqb += g * cf * vt * rhob;
        vtb = g * cf * sq * rhob - kb * bf / (2.0 * sq) * sqb;
        bfb = -kb * vt / (2.0 * sq) * sqb;
    }
    c2b += cfb / (2.0 * cf);
    dub = sgn * a.two_h * gunb;
    dvb = sgn * a.two_h * gvnb;
}

// The traces of side s of an element: du (normal derivative of u), vt, dv.
template <class D, int s>
__device__ __forceinline__ void traces3(cdbl* T, const double (&U)[D::NU * D::NU], const double (&V)[D::NV * D::NV],
                                        double (&du)[D::NQ], double (&vt)[D::NQ], double (&dv)[D::NQ]) {
    side_trace<D, D::NU, true, s>(T, D::oE1r, U, du);
    side_trace<D, D::NV, false, s>(T, D::oEr, V, vt);
    side_trace<D, D::NV, true, s>(T, D::oE1r, V, dv);
}

// ---------------------------------------------------------------------------
// k_adj_a: the element's own part of the VJP of one stage. The phases are
// ordered so that the two solves' results die early: lambda = M^-1 vbar_t
// feeds the beta and v terms of M, the lift adjoints of A and D and the v-test
// volume terms, then bbar = A^-1 ubar_t feeds the lift adjoints of F and the
// u-test volume term.

// the adjoints of slot q (0 A, 1 D: from lambda with the v rows; 2 F: from bbar
// with the u rows) of the four sides (src/rhs.jl:434-473), into the slot buffer
template <class D, int n, int q, class KA>
__device__ __forceinline__ void lift_adjoint(const KA& a, cdbl* T, int e, const double (&Xb)[n * n]) {
    constexpr int NQ = D::NQ;
    auto side = [&](auto sc) {
        constexpr int s = decltype(sc)::value;
        constexpr bool hz = s < 2, lo = (s == 0 || s == 2);
        constexpr int oR = q == 0 ? (lo ? D::oTEl : D::oTEr) : q == 1 ? (lo ? D::oFE1l : D::oFE1r)
                                                                    : (lo ? D::oE1l : D::oE1r);
        constexpr double sg = q == 2 ? -1.0 : 1.0;   // b = ... - Fu
        double Y[n];
#pragma unroll
        for (int x = 0; x < n; ++x) {
            double sa = 0.0;
#pragma unroll
            for (int y = 0; y < n; ++y) sa = fma(T[oR + y], hz ? Xb[y + n * x] : Xb[x + n * y], sa);
            Y[x] = sg * sa;
        }
        double S[NQ];
        rows_sym<NQ, n, false>(T + D::oP, [&](int b) { return Y[b]; }, S);
        double* sl = a.slot + (long)(s * 3 + q) * NQ * a.E + e;
#pragma unroll
        for (int i = 0; i < NQ; ++i) sl[(long)i * a.E] = T[D::oW + i] * S[i];
    };
    side(std::integral_constant<int, 0>());
    side(std::integral_constant<int, 1>());
    side(std::integral_constant<int, 2>());
    side(std::integral_constant<int, 3>());
}

template <int NU, int NV, int W = 0>
__device__ __forceinline__ void adj_a_body(int e, int lane, double* Gl) {
    using D = Dims<NU, NV>;
    constexpr int NQ = D::NQ, NH = D::NH, nU = NU * NU, nV = NV * NV;
    const int E = adjargs<W>().E;
    const double c2e = adjargs<W>().c2[e], ale = adjargs<W>().al[e], bee = adjargs<W>().be[e];
    // the v cotangent accumulates in the lane's LDS slots once the mass weights
    // are no longer needed (NH^2 >= NV^2 entries per lane); the u cotangent is
    // stored as soon as it is formed
    auto vacc = [&](int m) -> double& { return Gl[m * kWave + lane]; };
    double betab = 0.0, c2b = 0.0;
    {
        // ---- mass system: lambda = M^-1 vbar_t (src/rhs.jl:886-939)
        double lam[nV];
        {
            auto& a = adjargs<W>();
            cdbl* T = ctab(a.tab);
            double V[nV], R[nV];
            ld<nV>(a.yv, E, e, V);
            ld<nV>(a.kbv, E, e, R);
#pragma unroll
            for (int m = 0; m < nV; ++m) R[m] *= a.kscale;
            double th = 1.0, de = 1.0, s1 = 1.0;
            const int n = mass_setup<D>(T, bee, V, Gl, lane, a.cheb_L, th, de, s1);
            if (n == 0) {
                atomicOr(a.fail, 1);
#pragma unroll
                for (int m = 0; m < nV; ++m) lam[m] = 0.0;
            } else {
                cheb_solve<D>(T + D::oPh, Gl, lane, n, th, de, s1, R, lam);
            }
        }
        phase_barrier();
        // ---- the beta and v terms of M: beta_bar = sum 2 W v_q Lambda_q X_q,
        // vbar += K(G') X with G' = 2 beta W Lambda_q
        {
            auto& a = adjargs<W>();
            cdbl* T = ctab(a.tab);
            double V[nV], X[nV];
            ld<nV>(a.yv, E, e, V);
            ld<nV>(a.kv, E, e, X);
#pragma unroll 1
            for (int p = 0; p < (NH + 1) / 2; ++p) {
                const bool mid = (NH & 1) && p == NH / 2;
                double lp[NH], lq[NH], yp[NH], yq[NH], xp[NH], xq[NH];
                vq_rows<D>(T + D::oPh, p, mid, lam, lp, lq);
                vq_rows<D>(T + D::oPh, p, mid, V, yp, yq);
                vq_rows<D>(T + D::oPh, p, mid, X, xp, xq);
                const int r = NH - 1 - p;
#pragma unroll
                for (int q = 0; q < NH; ++q) {
                    const double w1 = 2.0 * T[D::oWh + p + NH * q];
                    betab = fma(w1 * yp[q], lp[q] * xp[q], betab);
                    Gl[(p * NH + q) * kWave + lane] = (w1 * bee) * lp[q];
                    if (!mid) {
                        const double w2 = 2.0 * T[D::oWh + r + NH * q];
                        betab = fma(w2 * yq[q], lq[q] * xq[q], betab);
                        Gl[(r * NH + q) * kWave + lane] = (w2 * bee) * lq[q];
                    }
                }
            }
            double Vb[nV];
#pragma unroll
            for (int m = 0; m < nV; ++m) Vb[m] = 0.0;
            kapply<D>(T + D::oPh, Gl, lane, X, Vb);
#pragma unroll
            for (int m = 0; m < nV; ++m) vacc(m) = Vb[m];
        }
        phase_barrier();
        lift_adjoint<D, NV, 0>(adjargs<W>(), ctab(adjargs<W>().tab), e, lam);
        lift_adjoint<D, NV, 1>(adjargs<W>(), ctab(adjargs<W>().tab), e, lam);
        phase_barrier();
        // ---- v-test volume terms (src/rhs.jl:317-333): Rvol = c2 Vol(U) + alpha Vol(V),
        // Rv = -(4/h^2) Rvol + ...: Ubar = c2 Vol^T(Rbar), c2bar += <Vol^T(Rbar), U>
        const double s4 = -adjargs<W>().four_hh;
#pragma unroll
        for (int m = 0; m < nV; ++m) lam[m] *= s4;
        {
            auto& a = adjargs<W>();
            double Z[nU];
#pragma unroll
            for (int m = 0; m < nU; ++m) Z[m] = 0.0;
            vol_pass<D, NV, NU>(a.tab, lam, a.atab, Z);
            double U[nU];
            ld<nU>(a.yu, E, e, U);
#pragma unroll
            for (int m = 0; m < nU; ++m) {
                c2b = fma(Z[m], U[m], c2b);
                Z[m] *= c2e;
            }
            st<nU>(a.pu, E, e, Z);
        }
        phase_barrier();
        {
            auto& a = adjargs<W>();
            double Zv[nV];
#pragma unroll
            for (int m = 0; m < nV; ++m) Zv[m] = 0.0;
            vol_pass<D, NV, NV>(a.tab, lam, a.atab, Zv);
#pragma unroll
            for (int m = 0; m < nV; ++m) vacc(m) = fma(ale, Zv[m], vacc(m));
        }
    }
    phase_barrier();
    {
        // ---- u-equation (src/rhs.jl:361-382): ut = A^-1 b; bbar = A^-1 ubar_t,
        // c2bar -= bbar^T K2 ut; b = VB - Fu + e0 V[0]
        double bb[nU];
        {
            auto& a = adjargs<W>();
            cdbl* T = ctab(a.tab);
            cdbl* AT = ctab(a.atab);
            double ub[nU], Ut[nU];
            ld<nU>(a.kbu, E, e, ub);
#pragma unroll
            for (int m = 0; m < nU; ++m) ub[m] *= a.kscale;
            usolve<D>(T + D::oQ, AT + NQ * NQ + NU * NU, c2e, ub, bb);
            ld<nU>(a.ku, E, e, Ut);
            c2b -= stiff_dot<D>(AT + NQ * NQ, bb, Ut);
        }
        vacc(0) += bb[0];
        phase_barrier();
        lift_adjoint<D, NU, 2>(adjargs<W>(), ctab(adjargs<W>().tab), e, bb);
        phase_barrier();
        auto& a = adjargs<W>();
        // VB = c2 Vol(V) (u-test): Vbar += c2 Vol^T(bbar), c2bar += <Vol^T(bbar), V>
        double Zv[nV];
#pragma unroll
        for (int m = 0; m < nV; ++m) Zv[m] = 0.0;
        vol_pass<D, NU, NV>(a.tab, bb, a.atab, Zv);
        double V[nV];
        ld<nV>(a.yv, E, e, V);
#pragma unroll
        for (int m = 0; m < nV; ++m) {
            const double z = Zv[m];
            c2b = fma(z, V[m], c2b);
            Zv[m] = fma(c2e, z, vacc(m));
        }
        st<nV>(a.pv, E, e, Zv);
    }
    auto& a = adjargs<W>();
    a.gc2[e] += c2b;
    a.gbe[e] += betab;
}

// ---------------------------------------------------------------------------
// k_adj_b: the faces of the element (its share of each face's adjoint) and the
// RK4 bookkeeping of the stage.
template <class D, int s, class KA>
__device__ __forceinline__ void adj_face(const KA& a, int e, int j, int k, const double (&U)[D::NU * D::NU],
                                         const double (&V)[D::NV * D::NV], double c2e, double ale, double bee,
                                         double (&Ub)[D::NU * D::NU], double (&Vb)[D::NV * D::NV], double& c2b,
                                         double& betab) {
    constexpr int NQ = D::NQ, NU = D::NU, NV = D::NV;
    constexpr bool plus = (s == 0 || s == 2);
    constexpr int ns = s ^ 1;
    const int E = a.E;
    int nb;
    bool bnd;
    side_nb(s, e, j, k, a.N, nb, bnd);
    cdbl* T = ctab(a.tab);
    double du[NQ], vt[NQ], dv[NQ];
    traces3<D, s>(T, U, V, du, vt, dv);
    double Ab[NQ], Db[NQ], Fb[NQ];
    {
        const double* sl = a.slotr + (long)(s * 3) * NQ * E + e;
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            Ab[i] = sl[(long)(0 * NQ + i) * E];
            Db[i] = sl[(long)(1 * NQ + i) * E];
            Fb[i] = sl[(long)(2 * NQ + i) * E];
        }
    }
    double dub[NQ], vtb[NQ], dvb[NQ];
    if (bnd) {
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            double c2x, bfx;
            bnd_slots_T(a, s, vt[i], du[i], dv[i], c2e, ale, bee, Ab[i], Db[i], Fb[i], vtb[i], dub[i], dvb[i], c2x, bfx);
            c2b += c2x;
            betab += bfx;
        }
    } else {
        double nU_[NU * NU], nV_[NV * NV];
        ld<NU * NU>(a.yu, E, nb, nU_);
        ld<NV * NV>(a.yv, E, nb, nV_);
        double ndu[NQ], nvt[NQ], ndv[NQ];
        traces3<D, ns>(T, nU_, nV_, ndu, nvt, ndv);
        const double c2n = a.c2[nb], aln = a.al[nb];
        const double* sl = a.slotr + (long)(ns * 3) * NQ * E + nb;
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            const double An = sl[(long)(0 * NQ + i) * E], Dn = sl[(long)(1 * NQ + i) * E], Fn = sl[(long)(2 * NQ + i) * E];
            double c2x;
            if (plus)
                face_slots_T(a, true, nvt[i], vt[i], ndu[i], du[i], ndv[i], dv[i], c2n, c2e, aln, ale, Ab[i], Db[i],
                             Fb[i], An, Dn, Fn, vtb[i], dub[i], dvb[i], c2x);
            else
                face_slots_T(a, false, vt[i], nvt[i], du[i], ndu[i], dv[i], ndv[i], c2e, c2n, ale, aln, An, Dn, Fn,
                             Ab[i], Db[i], Fb[i], vtb[i], dub[i], dvb[i], c2x);
            c2b += c2x;
        }
    }
    side_trace_T<D, NU, true, s>(T, D::oE1r, dub, Ub);
    side_trace_T<D, NV, false, s>(T, D::oEr, vtb, Vb);
    side_trace_T<D, NV, true, s>(T, D::oE1r, dvb, Vb);
}

template <int NU, int NV, int W = 0>
__device__ __forceinline__ void adj_b_body(int e) {
    using D = Dims<NU, NV>;
    constexpr int nU = NU * NU, nV = NV * NV;
    const int E = adjargs<W>().E, N = adjargs<W>().N, j = e % N, k = e / N;
    const double c2e = adjargs<W>().c2[e], ale = adjargs<W>().al[e], bee = adjargs<W>().be[e];
    double U[nU], V[nV], Ub[nU], Vb[nV];
    {
        auto& a = adjargs<W>();
        ld<nU>(a.yu, E, e, U);
        ld<nV>(a.yv, E, e, V);
        ld<nU>(a.pu, E, e, Ub);
        ld<nV>(a.pv, E, e, Vb);
    }
    double c2b = 0.0, betab = 0.0;
    adj_face<D, 0>(adjargs<W>(), e, j, k, U, V, c2e, ale, bee, Ub, Vb, c2b, betab);
    phase_barrier();
    adj_face<D, 1>(adjargs<W>(), e, j, k, U, V, c2e, ale, bee, Ub, Vb, c2b, betab);
    phase_barrier();
    adj_face<D, 2>(adjargs<W>(), e, j, k, U, V, c2e, ale, bee, Ub, Vb, c2b, betab);
    phase_barrier();
    adj_face<D, 3>(adjargs<W>(), e, j, k, U, V, c2e, ale, bee, Ub, Vb, c2b, betab);
    phase_barrier();
    auto& a = adjargs<W>();
    a.gc2[e] += c2b;
    a.gbe[e] += betab;
    // RK4 (src/solver.jl:289-321) backwards: ybar_n = ybar + sum_s Ybar_s; the
    // cotangent of k_{s-1} is ca ybar + cb Ybar_s (stage 0: one VJP, Ybar to acc)
    auto book = [&](auto& Yb, auto nn, double* yb, double* acc, double* nkb) {
        constexpr int n = decltype(nn)::value;
        if (a.stage == 0) {
            st<n>(acc, E, e, Yb);
            return;
        }
        double y[n], ac[n];
        if (a.stage == 1) {
            ld<n>(acc, E, e, ac);
#pragma unroll
            for (int m = 0; m < n; ++m) ac[m] += Yb[m];
            st<n>(yb, E, e, ac);
            return;
        }
        ld<n>(yb, E, e, y);
        if (a.stage == 4) {
#pragma unroll
            for (int m = 0; m < n; ++m) ac[m] = y[m] + Yb[m];
        } else {
            ld<n>(acc, E, e, ac);
#pragma unroll
            for (int m = 0; m < n; ++m) ac[m] += Yb[m];
        }
        st<n>(acc, E, e, ac);
#pragma unroll
        for (int m = 0; m < n; ++m) ac[m] = a.ca * y[m] + a.cb * Yb[m];
        st<n>(nkb, E, e, ac);
    };
    book(Ub, std::integral_constant<int, nU>(), a.ybu, a.accu, a.nkbu);
    book(Vb, std::integral_constant<int, nV>(), a.ybv, a.accv, a.nkbv);
}

// ---------------------------------------------------------------------------
// Tangent-linear right-hand side: dF = F'(Y; c2, beta)(dY, dc2, dbeta) with
// the forward derivative k = F(Y) given.
template <class D, int s, class KA>
__device__ __forceinline__ void tl_face(const KA& a, int e, int j, int k, const double (&U)[D::NU * D::NU],
                                        const double (&V)[D::NV * D::NV], const double (&dU)[D::NU * D::NU],
                                        const double (&dV)[D::NV * D::NV], double c2e, double ale, double bee,
                                        double tc2e, double tbee, double (&Rv)[D::NV * D::NV],
                                        double (&bv)[D::NU * D::NU]) {
    constexpr int NQ = D::NQ, NU = D::NU, NV = D::NV;
    constexpr bool plus = (s == 0 || s == 2), hz = s < 2;
    constexpr int ns = s ^ 1;
    const int E = a.E;
    int nb;
    bool bnd;
    side_nb(s, e, j, k, a.N, nb, bnd);
    cdbl* T = ctab(a.tab);
    double du[NQ], vt[NQ], dv[NQ], tdu[NQ], tvt[NQ], tdv[NQ];
    traces3<D, s>(T, U, V, du, vt, dv);
    traces3<D, s>(T, dU, dV, tdu, tvt, tdv);
    double A[NQ], Dd[NQ], F[NQ];
    if (bnd) {
#pragma unroll
        for (int i = 0; i < NQ; ++i)
            bnd_slots_tl(a, s, vt[i], du[i], dv[i], c2e, ale, bee, tvt[i], tdu[i], tdv[i], tc2e, tbee, A[i], Dd[i], F[i]);
    } else {
        double ndu[NQ], nvt[NQ], ndv[NQ], ntdu[NQ], ntvt[NQ], ntdv[NQ];
        {
            double X[NU * NU], Y[NV * NV];
            ld<NU * NU>(a.yu, E, nb, X);
            ld<NV * NV>(a.yv, E, nb, Y);
            traces3<D, ns>(T, X, Y, ndu, nvt, ndv);
            ld<NU * NU>(a.du, E, nb, X);
            ld<NV * NV>(a.dv, E, nb, Y);
            traces3<D, ns>(T, X, Y, ntdu, ntvt, ntdv);
        }
        const double c2n = a.c2[nb], aln = a.al[nb], tc2n = a.dc2[nb];
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            if (plus)
                face_slots_tl(a, true, nvt[i], vt[i], ndu[i], du[i], ndv[i], dv[i], c2n, c2e, aln, ale, ntvt[i],
                              tvt[i], ntdu[i], tdu[i], ntdv[i], tdv[i], tc2n, tc2e, A[i], Dd[i], F[i]);
            else
                face_slots_tl(a, false, vt[i], nvt[i], du[i], ndu[i], dv[i], ndv[i], c2e, c2n, ale, aln, tvt[i],
                              ntvt[i], tdu[i], ntdu[i], tdv[i], ntdv[i], tc2e, tc2n, A[i], Dd[i], F[i]);
        }
    }
    // lift (src/rhs.jl:434-473)
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        const double w = T[D::oW + i];
        A[i] *= w; Dd[i] *= w; F[i] *= w;
    }
    double YA[NV], YD[NV], YF[NU];
    cols_sym<NQ, NV, false>(T + D::oP, A, YA);
    cols_sym<NQ, NV, false>(T + D::oP, Dd, YD);
    cols_sym<NQ, NU, false>(T + D::oP, F, YF);
    constexpr bool lo = plus;
    constexpr int oE = lo ? D::oTEl : D::oTEr, oE1 = lo ? D::oFE1l : D::oFE1r, oF1 = lo ? D::oE1l : D::oE1r;
#pragma unroll
    for (int b = 0; b < NV; ++b)
#pragma unroll
        for (int x = 0; x < NV; ++x) {
            double r = Rv[x + NV * b];
            if (hz) { r = fma(T[oE + x], YA[b], r); r = fma(T[oE1 + x], YD[b], r); }
            else { r = fma(YA[x], T[oE + b], r); r = fma(YD[x], T[oE1 + b], r); }
            Rv[x + NV * b] = r;
        }
#pragma unroll
    for (int b = 0; b < NU; ++b)
#pragma unroll
        for (int x = 0; x < NU; ++x)
            bv[x + NU * b] = hz ? fma(-T[oF1 + x], YF[b], bv[x + NU * b]) : fma(-YF[x], T[oF1 + b], bv[x + NU * b]);
}

template <int NU, int NV>
__device__ __forceinline__ void tl_body(int e, int lane, double* Gl) {
    using D = Dims<NU, NV>;
    constexpr int NH = D::NH, nU = NU * NU, nV = NV * NV;
    auto& a = adjargs();
    const int E = a.E, N = a.N, j = e % N, k = e / N;
    cdbl* T = ctab(a.tab);
    cdbl* AT = ctab(a.atab);
    const double c2e = a.c2[e], ale = a.al[e], bee = a.be[e], tc2e = a.dc2[e], tbee = a.dbe[e];
    double Rv[nV], bv[nU];
    {
        double U[nU], V[nV], dU[nU], dV[nV];
        ld<nU>(a.yu, E, e, U);
        ld<nV>(a.yv, E, e, V);
        ld<nU>(a.du, E, e, dU);
        ld<nV>(a.dv, E, e, dV);
#pragma unroll
        for (int m = 0; m < nV; ++m) Rv[m] = 0.0;
#pragma unroll
        for (int m = 0; m < nU; ++m) bv[m] = 0.0;
        tl_face<D, 0>(a, e, j, k, U, V, dU, dV, c2e, ale, bee, tc2e, tbee, Rv, bv);
        tl_face<D, 1>(a, e, j, k, U, V, dU, dV, c2e, ale, bee, tc2e, tbee, Rv, bv);
        tl_face<D, 2>(a, e, j, k, U, V, dU, dV, c2e, ale, bee, tc2e, tbee, Rv, bv);
        tl_face<D, 3>(a, e, j, k, U, V, dU, dV, c2e, ale, bee, tc2e, tbee, Rv, bv);
        // volume terms: dRvol = Vol(c2 dU + dc2 U) + Vol(alpha dV), dVB = Vol(c2 dV + dc2 V)
        double Rvol[nV];
#pragma unroll
        for (int m = 0; m < nV; ++m) Rvol[m] = 0.0;
        double X[nU];
#pragma unroll
        for (int m = 0; m < nU; ++m) X[m] = c2e * dU[m] + tc2e * U[m];
        vol_pass<D, NU, NV>(a.tab, X, a.atab, Rvol);
        double Y[nV];
#pragma unroll
        for (int m = 0; m < nV; ++m) Y[m] = ale * dV[m];
        vol_pass<D, NV, NV>(a.tab, Y, a.atab, Rvol);
#pragma unroll
        for (int m = 0; m < nV; ++m) Y[m] = c2e * dV[m] + tc2e * V[m];
        vol_pass<D, NV, NU>(a.tab, Y, a.atab, bv);
#pragma unroll
        for (int m = 0; m < nV; ++m) Rv[m] = fma(-a.four_hh, Rvol[m], Rv[m]);
        bv[0] += dV[0];
    }
    // u-equation: d ut = A^-1 (db - dc2 K2 ut)
    {
        double Ut[nU], KU[nU];
        ld<nU>(a.ku, E, e, Ut);
        cdbl* K = AT + D::NQ * D::NQ;
#pragma unroll
        for (int bb = 0; bb < NU; ++bb)
#pragma unroll
            for (int x = 0; x < NU; ++x) {
                double t = 0.0;
#pragma unroll
                for (int c = 0; c < NU; ++c) {
                    t = fma(K[x + NU * c], Ut[c + NU * bb], t);
                    t = fma(Ut[x + NU * c], K[c + NU * bb], t);
                }
                KU[x + NU * bb] = bv[x + NU * bb] - tc2e * t;
            }
        double o[nU];
        usolve<D>(T + D::oQ, AT + D::NQ * D::NQ + NU * NU, c2e, KU, o);
        st<nU>(a.outu, E, e, o);
    }
    // mass: dX = M^-1 (dRv + dK X), dK with dG = 2 W (dbeta v_q + beta dv_q)
    {
        double V[nV], dV[nV], X[nV];
        ld<nV>(a.yv, E, e, V);
        ld<nV>(a.dv, E, e, dV);
        ld<nV>(a.kv, E, e, X);
#pragma unroll
        for (int p = 0; p < (NH + 1) / 2; ++p) {
            const bool mid = (NH & 1) && p == NH / 2;
            double vp_[NH], vq_[NH], tp[NH], tq[NH];
            vq_rows<D>(T + D::oPh, p, mid, V, vp_, vq_);
            vq_rows<D>(T + D::oPh, p, mid, dV, tp, tq);
#pragma unroll
            for (int q = 0; q < NH; ++q) {
                Gl[(p * NH + q) * kWave + lane] = 2.0 * T[D::oWh + p + NH * q] * (tbee * vp_[q] + bee * tp[q]);
                if (!mid) {
                    const int r = NH - 1 - p;
                    Gl[(r * NH + q) * kWave + lane] = 2.0 * T[D::oWh + r + NH * q] * (tbee * vq_[q] + bee * tq[q]);
                }
            }
        }
        kapply<D>(T + D::oPh, Gl, lane, X, Rv);
        double th = 1.0, de = 1.0, s1 = 1.0, o[nV];
        const int n = mass_setup<D>(T, bee, V, Gl, lane, a.cheb_L, th, de, s1);
        if (n == 0) {
            atomicOr(a.fail, 1);
#pragma unroll
            for (int m = 0; m < nV; ++m) o[m] = 0.0;
        } else {
            cheb_solve<D>(T + D::oPh, Gl, lane, n, th, de, s1, Rv, o);
        }
        st<nV>(a.outv, E, e, o);
    }
}

// ---------------------------------------------------------------------------
// Misfit of u_h on the bottom side (y = 0) at the face points of the elements
// of row j = 0 (e = N k): the trace, 1/2 (h/2) w_i (u_h - g)^2 per point, and
// with ybu set the misfit's derivative added to the state cotangent.
template <int NU, int W = 0>
__device__ __forceinline__ void misfit_body(int k) {
    auto& a = adjargs<W>();
    using D = Dims<NU, NU>;
    constexpr int NQ = D::NQ, nU = NU * NU;
    const int E = a.E, e = a.N * k;
    cdbl* T = ctab(a.tab);
    double U[nU], o[NQ];
    ld<nU>(a.yu, E, e, U);
    side_trace<D, NU, false, 0>(T, D::oEr, U, o);
    double r[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        if (a.tr) a.tr[(long)k * NQ + i] = o[i];
        const double g = a.tgt ? a.tgt[(long)k * NQ + i] : 0.0;
        const double w = a.hw * T[D::oW + i];
        r[i] = w * (o[i] - g);
        if (a.jp) a.jp[(long)k * NQ + i] = 0.5 * (o[i] - g) * r[i];
    }
    if (a.ybu && a.tgt) {
        double Y[nU];
        ld<nU>(a.ybu, E, e, Y);
        side_trace_T<D, NU, false, 0>(T, D::oEr, r, Y);
        st<nU>(a.ybu, E, e, Y);
    }
}

// ---------------------------------------------------------------------------
// Kernels. The element of a lane comes from block_first_element (XCD-aware
// order); lanes past E return (no kernel uses a block barrier).

template <int NU, int NV>
__global__ __launch_bounds__(kWave) void k_adj_a(const AdjArgs a) {
    using D = Dims<NU, NV>;
    __shared__ double Gl[D::NH * D::NH * kWave];
    (void)a;
    const int e = block_first_element<false>() + (int)threadIdx.x;
    if (e >= adjargs().E) return;
    adj_a_body<NU, NV>(e, threadIdx.x, Gl);
}

template <int NU, int NV>
__global__ __launch_bounds__(kWave) void k_adj_b(const AdjArgs a) {
    (void)a;
    const int e = block_first_element<false>() + (int)threadIdx.x;
    if (e >= adjargs().E) return;
    adj_b_body<NU, NV>(e);
}

// k_adj_b of stage s, then k_adj_a of stage s - 1 for the same element: the
// second part needs only the element's own cotangent from the first (the
// neighbours' slot adjoints of stage s are read from slotr while stage s - 1's
// are written to slot, a different buffer).
template <int NU, int NV>
__global__ __launch_bounds__(kWave) void k_adj_ba(const AdjArgs b, const AdjArgs a) {
    using D = Dims<NU, NV>;
    __shared__ double Gl[D::NH * D::NH * kWave];
    (void)a; (void)b;
    const int e = block_first_element<false>() + (int)threadIdx.x;
    if (e >= adjargs<0>().E) return;
    adj_b_body<NU, NV, 0>(e);
    phase_barrier();
    adj_a_body<NU, NV, 1>(e, threadIdx.x, Gl);
}

template <int NU, int NV>
__global__ __launch_bounds__(kWave) void k_tl(const AdjArgs a) {
    using D = Dims<NU, NV>;
    __shared__ double Gl[D::NH * D::NH * kWave];
    (void)a;
    const int e = block_first_element<false>() + (int)threadIdx.x;
    if (e >= adjargs().E) return;
    tl_body<NU, NV>(e, threadIdx.x, Gl);
}

template <int NU>
__global__ __launch_bounds__(kWave) void k_misfit(const AdjArgs a) {
    (void)a;
    const int k = blockIdx.x * kWave + (int)threadIdx.x;
    if (k >= adjargs().N) return;
    misfit_body<NU>(k);
}

// Batched forms: each block's instance and its place in the instance's range
// come from the BatchArgs block map; args holds an AdjArgs pair per instance
// (k_adj_a, k_adj_b and k_misfit use the first, k_adj_ba both).
template <int NU, int NV>
__global__ __launch_bounds__(kWave) void k_adj_a_bt(const BatchArgs b) {
    using D = Dims<NU, NV>;
    __shared__ double Gl[D::NH * D::NH * kWave];
    (void)b;
    const int e = block_first_element<true>() + (int)threadIdx.x;
    if (e >= adjargs<2>().E) return;
    adj_a_body<NU, NV, 2>(e, threadIdx.x, Gl);
}

template <int NU, int NV>
__global__ __launch_bounds__(kWave) void k_adj_b_bt(const BatchArgs b) {
    (void)b;
    const int e = block_first_element<true>() + (int)threadIdx.x;
    if (e >= adjargs<2>().E) return;
    adj_b_body<NU, NV, 2>(e);
}

template <int NU, int NV>
__global__ __launch_bounds__(kWave) void k_adj_ba_bt(const BatchArgs b) {
    using D = Dims<NU, NV>;
    __shared__ double Gl[D::NH * D::NH * kWave];
    (void)b;
    const int e = block_first_element<true>() + (int)threadIdx.x;
    if (e >= adjargs<2>().E) return;
    adj_b_body<NU, NV, 2>(e);
    phase_barrier();
    adj_a_body<NU, NV, 3>(e, threadIdx.x, Gl);
}

template <int NU>
__global__ __launch_bounds__(kWave) void k_misfit_bt(const BatchArgs b) {
    (void)b;
    const int k = block_first_element<true>() + (int)threadIdx.x;
    if (k >= adjargs<2>().N) return;
    misfit_body<NU, 2>(k);
}

// batched launches of one (n_u, n_v) pair over `grid` blocks; which: 0 k_adj_a,
// 1 k_adj_b, 2 k_adj_ba, 3 k_misfit (blocks over the bottom row's elements)
template <int NU, int NV>
static hipError_t launch_adjb_nv(const BatchArgs& b, int grid, int which, hipStream_t s) {
    switch (which) {
        case 0: k_adj_a_bt<NU, NV><<<grid, kWave, 0, s>>>(b); break;
        case 1: k_adj_b_bt<NU, NV><<<grid, kWave, 0, s>>>(b); break;
        case 2: k_adj_ba_bt<NU, NV><<<grid, kWave, 0, s>>>(b); break;
        case 3: k_misfit_bt<NU><<<grid, kWave, 0, s>>>(b); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// launches of one (n_u, n_v) pair; which: 0 k_adj_a, 1 k_adj_b, 2 k_tl, 3 k_misfit
template <int NU, int NV>
static hipError_t launch_adj_nv(const AdjArgs& a, int which, hipStream_t s) {
    const int grid = (a.E + kWave - 1) / kWave;
    switch (which) {
        case 0: k_adj_a<NU, NV><<<grid, kWave, 0, s>>>(a); break;
        case 1: k_adj_b<NU, NV><<<grid, kWave, 0, s>>>(a); break;
        case 2: k_tl<NU, NV><<<grid, kWave, 0, s>>>(a); break;
        case 3: k_misfit<NU><<<(a.N + kWave - 1) / kWave, kWave, 0, s>>>(a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int NU, int NV>
static hipError_t launch_adj2_nv(const AdjArgs& b, const AdjArgs& a, hipStream_t s) {
    k_adj_ba<NU, NV><<<(a.E + kWave - 1) / kWave, kWave, 0, s>>>(b, a);
    return hipGetLastError();
}

hipError_t launch_adj_q1(const AdjArgs& a, int nv, int which, hipStream_t s);
hipError_t launch_adj2_q1(const AdjArgs& b, const AdjArgs& a, int nv, hipStream_t s);
hipError_t launch_adj_q2(const AdjArgs& a, int nv, int which, hipStream_t s);
hipError_t launch_adj2_q2(const AdjArgs& b, const AdjArgs& a, int nv, hipStream_t s);
hipError_t launch_adj_q3(const AdjArgs& a, int nv, int which, hipStream_t s);
hipError_t launch_adj2_q3(const AdjArgs& b, const AdjArgs& a, int nv, hipStream_t s);
hipError_t launch_adj_q4(const AdjArgs& a, int nv, int which, hipStream_t s);
hipError_t launch_adj2_q4(const AdjArgs& b, const AdjArgs& a, int nv, hipStream_t s);
hipError_t launch_adj_q5(const AdjArgs& a, int nv, int which, hipStream_t s);
hipError_t launch_adj2_q5(const AdjArgs& b, const AdjArgs& a, int nv, hipStream_t s);
hipError_t launch_adj_q6(const AdjArgs& a, int nv, int which, hipStream_t s);
hipError_t launch_adj2_q6(const AdjArgs& b, const AdjArgs& a, int nv, hipStream_t s);
hipError_t launch_adjb_q1(const BatchArgs& b, int grid, int nv, int which, hipStream_t s);
hipError_t launch_adjb_q2(const BatchArgs& b, int grid, int nv, int which, hipStream_t s);
hipError_t launch_adjb_q3(const BatchArgs& b, int grid, int nv, int which, hipStream_t s);
hipError_t launch_adjb_q4(const BatchArgs& b, int grid, int nv, int which, hipStream_t s);
hipError_t launch_adjb_q5(const BatchArgs& b, int grid, int nv, int which, hipStream_t s);
hipError_t launch_adjb_q6(const BatchArgs& b, int grid, int nv, int which, hipStream_t s);

}  // namespace wedg
