/* CPU oracle: plain-C FP64 restatement of the reference's batched right-hand
 * side (src/rhs.jl:1-951) and classic RK4 (src/solver.jl:278-321).
 *
 * Test infrastructure only (tests/, the smoke check and the CPU-baseline leg of
 * bench.py use it as the checker). Built with -ffp-contract=off so that a*b+c is
 * never fused except where the reference writes muladd (traces, lifts, the
 * u-equation), which is fma() here. The element loops run in parallel with
 * OpenMP over contiguous element blocks, like the reference's task blocks
 * (src/rhs.jl:42-61); every element does the same operations in any block.
 *
 * Parity of each part is pinned by the tests (see oracle/README.md): against
 * the element-by-element dense restatement of test/reference_rhs.jl, the
 * initialization known answers, closed forms of the boundary fluxes, dense
 * mass solves, and the golden errors in data/convergence/results.csv.
 */
#define _GNU_SOURCE
#include "wedg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXN 8      /* q_u <= 7 */
#define MAXH 12     /* q_high <= 11 */

typedef struct {
    double *duB, *duT, *duL, *duR, *vB, *vT, *vL, *vR, *dvB, *dvT, *dvL, *dvR;
    double *AB, *AT, *AL, *AR, *DB, *DT, *DL, *DR, *FB, *FT, *FL, *FR;
    double *G, *th, *de, *s1;   /* Chebyshev interval data per element */
    int *ne;                     /* per-element step counts */
    double *buf;
} orc_ws;

static int ws_alloc(const orc_pre* p, orc_ws* w) {
    long E = (long)p->N * p->N, nq = p->qu1, qh = p->qh;
    long tot = 24 * nq * E + qh * qh * E + 3 * E;
    w->buf = (double*)calloc((size_t)tot, sizeof(double));
    w->ne = (int*)calloc((size_t)E, sizeof(int));
    if (!w->buf || !w->ne) { free(w->buf); free(w->ne); return -1; }
    double* q = w->buf;
    double** arr[24] = {&w->duB, &w->duT, &w->duL, &w->duR, &w->vB, &w->vT, &w->vL, &w->vR,
                        &w->dvB, &w->dvT, &w->dvL, &w->dvR, &w->AB, &w->AT, &w->AL, &w->AR,
                        &w->DB, &w->DT, &w->DL, &w->DR, &w->FB, &w->FT, &w->FL, &w->FR};
    for (int i = 0; i < 24; ++i) { *arr[i] = q; q += nq * E; }
    w->G = q; q += qh * qh * E;
    w->th = q; q += E;
    w->de = q; q += E;
    w->s1 = q;
    return 0;
}

static void ws_free(orc_ws* w) { free(w->buf); free(w->ne); }

/* ---- traces (src/rhs.jl:384-410, 573-580) ---- */
static void trace4(double* oB, double* oT, double* oL, double* oR, const double* P, int nq, int n,
                   const double* El, const double* Er, const double* X /* X[a + n b] */, long e) {
    long o = (long)nq * e;
    for (int i = 0; i < nq; ++i) { oB[o + i] = 0.0; oT[o + i] = 0.0; oL[o + i] = 0.0; oR[o + i] = 0.0; }
    for (int b = 0; b < n; ++b) {
        double sB = 0.0, sT = 0.0, sL = 0.0, sR = 0.0;
        for (int a = 0; a < n; ++a) {
            double x = X[a + n * b], xt = X[b + n * a];
            sB = fma(El[a], x, sB); sT = fma(Er[a], x, sT);
            sL = fma(El[a], xt, sL); sR = fma(Er[a], xt, sR);
        }
        for (int i = 0; i < nq; ++i) {
            double pp = P[i + nq * b];
            oB[o + i] = fma(pp, sB, oB[o + i]); oT[o + i] = fma(pp, sT, oT[o + i]);
            oL[o + i] = fma(pp, sL, oL[o + i]); oR[o + i] = fma(pp, sR, oR[o + i]);
        }
    }
}

static void load_elem(const double* x, long E, int n, long e, double* X) {
    for (int b = 0; b < n; ++b)
        for (int a = 0; a < n; ++a) X[a + n * b] = x[e + E * (a + (long)n * b)];
}

/* out (m x m) = L Z R^T with L, R (m x n) column-major, Z (n x n): T = L Z first */
static void cfwd(double* out, const double* L, const double* R, int m, int n, const double* Z) {
    double T[MAXH * MAXH];
    for (int b = 0; b < n; ++b)
        for (int i = 0; i < m; ++i) {
            double s = 0.0;
            for (int a = 0; a < n; ++a) s += L[i + m * a] * Z[a + n * b];
            T[i + m * b] = s;
        }
    for (int q = 0; q < m; ++q)
        for (int i = 0; i < m; ++i) {
            double s = 0.0;
            for (int b = 0; b < n; ++b) s += T[i + m * b] * R[q + m * b];
            out[i + m * q] = s;
        }
}

/* out (n x n) = L^T H R (+ out if acc) with L, R (m x n), H (m x m): S = L^T H first */
static void cbwd(double* out, const double* L, const double* R, int m, int n, const double* H, int acc) {
    double S[MAXH * MAXH];
    for (int q = 0; q < m; ++q)
        for (int a = 0; a < n; ++a) {
            double s = 0.0;
            for (int pp = 0; pp < m; ++pp) s += L[pp + m * a] * H[pp + m * q];
            S[a + n * q] = s;
        }
    for (int b = 0; b < n; ++b)
        for (int a = 0; a < n; ++a) {
            double s = 0.0;
            for (int q = 0; q < m; ++q) s += S[a + n * q] * R[q + m * b];
            out[a + n * b] = acc ? s + out[a + n * b] : s;
        }
}

/* P (m x n) restricted from a column-major table with leading dimension m */
static int acosh_steps(double sigma1, double mass_tol) {
    double L = acosh(1.0 / mass_tol);
    return (int)ceil(L / acosh(sigma1));
}

/* Chebyshev interval of one element (src/rhs.jl:889-907). Returns the step
 * count, or -1 if 1 - 2 beta v_q <= 0 somewhere. Fills G = 2 beta v_q W. */
static int cheb_setup_elem(const orc_pre* p, const double* V, long e, double* G, double* th,
                           double* de, double* s1) {
    int nv = p->qv1, qh = p->qh;
    long E = (long)p->N * p->N;
    cfwd(G, p->P_h_v, p->P_h_v, qh, nv, V);
    double gmin = INFINITY, gmax = -INFINITY;
    for (int iq = 0; iq < qh; ++iq)
        for (int ip = 0; ip < qh; ++ip) {
            double b = p->beta_hq[ip + (long)qh * (e + E * iq)];
            double g = G[ip + qh * iq] * (2.0 * b);
            G[ip + qh * iq] = g;
            if (g < gmin) gmin = g;
            if (g > gmax) gmax = g;
        }
    double amin = 1.0 - gmax, amax = 1.0 - gmin;
    if (!(amin > 0)) return -1;
    double t = 0.5 * (amax + amin);
    double d = fmax(0.5 * (amax - amin), 1e-15 * t);
    *th = t; *de = d; *s1 = t / d;
    int n = acosh_steps(*s1, p->mass_tol);
    if (n < 1) n = 1;
    for (int iq = 0; iq < qh; ++iq)
        for (int ip = 0; ip < qh; ++ip) G[ip + qh * iq] *= p->W2d_h[ip + qh * iq];
    return n;
}

/* n Chebyshev steps on one element (src/rhs.jl:921-939) */
static void cheb_iter_elem(const orc_pre* p, const double* G, double th, double de, double s1,
                           int n, const double* R, double* X) {
    int nv = p->qv1, qh = p->qh, n2 = nv * nv;
    double r[MAXN * MAXN], d[MAXN * MAXN], Kd[MAXN * MAXN], Y[MAXH * MAXH];
    for (int i = 0; i < n2; ++i) { r[i] = R[i]; d[i] = r[i] / th; X[i] = 0.0; }
    double rho = 1.0 / s1;
    for (int step = 1; step <= n; ++step) {
        for (int i = 0; i < n2; ++i) X[i] += d[i];
        if (step == n) break;
        cfwd(Y, p->P_h_v, p->P_h_v, qh, nv, d);
        for (int i = 0; i < qh * qh; ++i) Y[i] *= G[i];
        cbwd(Kd, p->P_h_v, p->P_h_v, qh, nv, Y, 0);
        for (int i = 0; i < n2; ++i) r[i] -= d[i] - Kd[i];
        double rhon = 1.0 / (2.0 * s1 - rho);
        double c1 = rhon * rho, c2 = 2.0 * rhon / de;
        for (int i = 0; i < n2; ++i) d[i] = c1 * d[i] + c2 * r[i];
        rho = rhon;
    }
}

/* ---- conjugate gradients on one element (src/rhs.jl:795-853) ---- */
static void mul_loops(double* C, const double* A, const double* B, int m, int l, int n) {
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            double s = 0.0;
            for (int k = 0; k < l; ++k) s = fma(A[i + m * k], B[k + l * j], s);
            C[i + m * j] = s;
        }
}

static double dotn(const double* a, const double* b, int n) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static void applyK(const orc_pre* p, const double* Pt, const double* G, const double* Z, double* out) {
    int nv = p->qv1, qh = p->qh;
    double T1[MAXH * MAXN], Y[MAXH * MAXH], T2[MAXN * MAXH];
    mul_loops(T1, p->P_h_v, Z, qh, nv, nv);
    mul_loops(Y, T1, Pt, qh, nv, qh);
    for (int i = 0; i < qh * qh; ++i) Y[i] *= G[i];
    mul_loops(T2, Pt, Y, nv, qh, qh);
    mul_loops(out, T2, p->P_h_v, nv, qh, nv);
}

/* returns 0 converged, 1 not converged, -1 not positive definite */
static int cg_elem(const orc_pre* p, const double* V, const double* R, double* X, long e) {
    int nv = p->qv1, qh = p->qh, n2 = nv * nv;
    int N = p->N;
    long E = (long)N * N;
    int j = (int)(e % N), k = (int)(e / N);
    double Pt[MAXN * MAXH], T1[MAXH * MAXN], Y[MAXH * MAXH], G[MAXH * MAXH];
    double r[MAXN * MAXN], pp[MAXN * MAXN], Ap[MAXN * MAXN];
    for (int a = 0; a < nv; ++a)
        for (int i = 0; i < qh; ++i) Pt[a + nv * i] = p->P_h_v[i + qh * a];
    mul_loops(T1, p->P_h_v, V, qh, nv, nv);
    mul_loops(Y, T1, Pt, qh, nv, qh);
    for (int iq = 0; iq < qh; ++iq)
        for (int ip = 0; ip < qh; ++ip) {
            /* beta_hE (qh, qh, N, N) */
            double b = p->beta_hE[ip + qh * (iq + (long)qh * (j + (long)N * k))];
            G[ip + qh * iq] = 2.0 * b * Y[ip + qh * iq] * p->W2d_h[ip + qh * iq];
        }
    (void)E;
    for (int i = 0; i < n2; ++i) X[i] = R[i];
    applyK(p, Pt, G, X, r);
    for (int i = 0; i < n2; ++i) pp[i] = r[i];
    double rr = dotn(r, r, n2), r0 = sqrt(dotn(R, R, n2));
    int it = 0, maxit = n2 + 10;
    while (sqrt(rr) > p->mass_tol * r0 && it < maxit) {
        ++it;
        applyK(p, Pt, G, pp, Ap);
        for (int i = 0; i < n2; ++i) Ap[i] = pp[i] - Ap[i];
        double pAp = dotn(pp, Ap, n2);
        if (!(pAp > 0)) return -1;
        double a = rr / pAp;
        for (int i = 0; i < n2; ++i) { X[i] += a * pp[i]; r[i] -= a * Ap[i]; }
        double rrn = dotn(r, r, n2);
        double bb = rrn / rr;
        for (int i = 0; i < n2; ++i) pp[i] = r[i] + bb * pp[i];
        rr = rrn;
    }
    return sqrt(rr) > p->mass_tol * r0 ? 1 : 0;
}

/* ---- boundary fluxes (src/rhs.jl:74-116) ---- */
static void side_flux(const orc_pre* p, int s, double c2f, double af, double cf, double bf, double v_fp,
                      double gun, double gvn, double gu_ex, double gv_ex, double v_ex, double* F1, double* dv) {
    if (p->kind[s] == 3) {
        double kap = p->kap[s];
        double g2 = c2f * gu_ex + af * gv_ex;
        double rho_w = c2f * gun + af * gvn - g2;
        *F1 = g2;
        *dv = -(kap / cf) * rho_w;
        return;
    }
    double g = p->gam[s], et = p->eta[s], sg = p->sig[s], kb = p->kb[s];
    double rho_num = et * (c2f * gun + af * gvn) + g * cf * sqrt(1.0 - kb * bf * v_fp) * v_fp;
    double rho_ex = et * (c2f * gu_ex + af * gv_ex) + g * cf * sqrt(1.0 - kb * bf * v_ex) * v_ex;
    double rho_net = rho_num - rho_ex;
    double sqf = sqrt(1.0 - kb * bf * v_fp);
    *F1 = c2f * gun + af * gvn - (et + sg * g) * rho_net;
    *dv = -(g - sg * et) / (cf * sqf) * rho_net;
}

static int weight_check(double kb, double bf, double v_fp, double v_ex) {
    return 1.0 - kb * bf * v_fp <= 0.0 ? 1 : (1.0 - kb * bf * v_ex <= 0.0 ? 2 : 0);
}

/* one side, point by point (src/rhs.jl:738-765); returns the smallest key or 0 */
static long boundary_side(const orc_pre* p, orc_ws* w, int s, double ct, double st, double two_h) {
    int N = p->N, nq = p->qu1;
    int hz = s < 2;
    double* A = s == 0 ? w->AB : s == 1 ? w->AT : s == 2 ? w->AL : w->AR;
    double* D = s == 0 ? w->DB : s == 1 ? w->DT : s == 2 ? w->DL : w->DR;
    double* Fs = s == 0 ? w->FB : s == 1 ? w->FT : s == 2 ? w->FL : w->FR;
    const double* vtr = s == 0 ? w->vB : s == 1 ? w->vT : s == 2 ? w->vL : w->vR;
    const double* dutr = s == 0 ? w->duB : s == 1 ? w->duT : s == 2 ? w->duL : w->duR;
    const double* dvtr = s == 0 ? w->dvB : s == 1 ? w->dvT : s == 2 ? w->dvL : w->dvR;
    const double* c2a = hz ? p->hc2 : p->vc2;
    const double* afa = hz ? p->haf : p->vaf;
    const double* cfa = hz ? p->hcf : p->vcf;
    const double* bfa = hz ? p->hbf : p->vbf;
    const double* sp = s == 0 ? p->bcB_sp : s == 1 ? p->bcT_sp : s == 2 ? p->bcL_sp : p->bcR_sp;
    const double* spn = s == 0 ? p->bcB_spy : s == 1 ? p->bcT_spy : s == 2 ? p->bcL_spx : p->bcR_spx;
    int rc = (s == 0 || s == 2) ? 0 : N - 1;     /* element row (hz) / column (vt) */
    int fc = (s == 0 || s == 2) ? 0 : N;         /* face row / column */
    double sgn = (s == 0 || s == 2) ? -1.0 : 1.0;
    int wall = p->kind[s] == 3;
    long key = 0;
    for (int m = 0; m < N; ++m)
        for (int i = 0; i < nq; ++i) {
            long I, If;
            if (hz) { I = i + (long)nq * (rc + (long)N * m); If = i + (long)nq * (fc + (long)(N + 1) * m); }
            else    { I = i + (long)nq * (m + (long)N * rc); If = i + (long)nq * (m + (long)N * fc); }
            double c2f = c2a[If], af = afa[If], cf = cfa[If], bf = bfa[If];
            double v_fp = vtr[I];
            double gun = sgn * (two_h * dutr[I]);
            double gvn = sgn * (two_h * dvtr[I]);
            double spv = sp[i + (long)nq * m], spnv = spn[i + (long)nq * m];
            double gu_ex = sgn * (ct * spnv - spnv);
            double gv_ex = sgn * (-p->omega * st * spnv);
            double v_ex = -p->omega * st * spv;
            if (!wall) {
                int check = weight_check(p->kb[s], bf, v_fp, v_ex);
                if (check) {
                    long k2 = 2 * ((long)s * nq * N + i + (long)m * nq) + check;
                    key = key == 0 ? k2 : (k2 < key ? k2 : key);
                    continue;
                }
            }
            double F1, dv;
            side_flux(p, s, c2f, af, cf, bf, v_fp, gun, gvn, gu_ex, gv_ex, v_ex, &F1, &dv);
            A[I] = F1;
            D[I] = sgn * (af * dv);
            Fs[I] = -sgn * (c2f * dv);
        }
    return key;
}

/* ---- interior faces (src/rhs.jl:620-700) ----
 * hz: "+" element (j, k), "-" element (j-1, k), face (i, j, k) of hface arrays.
 * vt: "+" element (j, k), "-" element (j, k-1), face (i, j, k) of vface arrays. */
static void interior_faces(const orc_pre* p, orc_ws* w, int hz, double two_h) {
    int N = p->N, nq = p->qu1;
    double th = p->theta, tau = p->tau, lu = p->lambda_u, lv = p->lambda_v;
    double *Ap = hz ? w->AB : w->AL, *Dp = hz ? w->DB : w->DL, *Fp = hz ? w->FB : w->FL;
    double *Am = hz ? w->AT : w->AR, *Dm = hz ? w->DT : w->DR, *Fm = hz ? w->FT : w->FR;
    const double *vM = hz ? w->vT : w->vR, *vP = hz ? w->vB : w->vL;
    const double *duM = hz ? w->duT : w->duR, *duP = hz ? w->duB : w->duL;
    const double *dvM = hz ? w->dvT : w->dvR, *dvP = hz ? w->dvB : w->dvL;
    if (N < 2) return;
#pragma omp parallel for schedule(static)
    for (int k = (hz ? 0 : 1); k < N; ++k)
        for (int j = (hz ? 1 : 0); j < N; ++j) {
            int jm = hz ? j - 1 : j, km = hz ? k : k - 1;
            long ep = j + (long)N * k, em = jm + (long)N * km;
            for (int i = 0; i < nq; ++i) {
                long fi = hz ? i + (long)nq * (j + (long)(N + 1) * k) : i + (long)nq * (j + (long)N * k);
                double vm = vM[i + nq * em], vp = vP[i + nq * ep];
                double dum = duM[i + nq * em], dup = duP[i + nq * ep];
                double dvm = dvM[i + nq * em], dvp = dvP[i + nq * ep];
                if (p->one_sided) {
                    double c2m = (hz ? p->hc2m : p->vc2m)[fi], c2p = (hz ? p->hc2 : p->vc2)[fi];
                    double afm = (hz ? p->hafm : p->vafm)[fi], afp = (hz ? p->haf : p->vaf)[fi];
                    double K = (hz ? p->hJC : p->vJC)[fi];
                    double J = two_h * (c2m * dum - c2p * dup + afm * dvm - afp * dvp) + K;
                    double G = two_h * ((1 - th) * (c2m * dum + afm * dvm) + th * (c2p * dup + afp * dvp));
                    double pen = (0.5 * (c2m + c2p)) * lu + (0.5 * (afm + afp)) * lv;
                    double s = vm - vp;
                    double dvsp = th * s - tau * J;
                    double dvsm = (1 - th) * (vp - vm) - tau * J;
                    Ap[i + nq * ep] = pen * s - G - (1 - th) * K;
                    Dp[i + nq * ep] = -afp * dvsp;
                    Fp[i + nq * ep] = c2p * dvsp;
                    Am[i + nq * em] = G - pen * s - th * K;
                    Dm[i + nq * em] = afm * dvsm;
                    Fm[i + nq * em] = -c2m * dvsm;
                } else {
                    double c2 = (hz ? p->hc2 : p->vc2)[fi], af = (hz ? p->haf : p->vaf)[fi];
                    double jmp = vm - vp;
                    double gst = two_h * (c2 * (th * dup + (1 - th) * dum) + af * (th * dvp + (1 - th) * dvm));
                    double tj = tau * (c2 * (two_h * (dum - dup)) + af * (two_h * (dvm - dvp)));
                    Ap[i + nq * ep] = (c2 * lu + af * lv) * jmp - gst;
                    Dp[i + nq * ep] = -af * (th * jmp - tj);
                    Fp[i + nq * ep] = c2 * (th * jmp - tj);
                    Am[i + nq * em] = gst - (c2 * lu + af * lv) * jmp;
                    Dm[i + nq * em] = af * ((1 - th) * (vp - vm) - tj);
                    Fm[i + nq * em] = -c2 * ((1 - th) * (vp - vm) - tj);
                }
            }
        }
}

/* ---- region 3 on one element: volume, source, lifts, u-equation ---- */
static void element_rhs(const orc_pre* p, const orc_ws* w, long e, const double* U, const double* V,
                        double ct, double st, double srct, double two_h, double four_hh,
                        double* Rv, double* UT) {
    int nu = p->qu1, nv = p->qv1, nq = p->qu1, N = p->N;
    long E = (long)N * N;
    double G1[MAXN * MAXN], G2[MAXN * MAXN], H1[MAXN * MAXN], H2[MAXN * MAXN];
    double VB[MAXN * MAXN], Fu[MAXN * MAXN], Cw[MAXN * MAXN], Aw[MAXN * MAXN];
    for (int ix = 0; ix < nq; ++ix)
        for (int iy = 0; iy < nq; ++iy) {
            Cw[iy + nq * ix] = p->Cw[iy + (long)nq * (e + E * ix)];
            Aw[iy + nq * ix] = p->Aw[iy + (long)nq * (e + E * ix)];
        }
    const double *P_u = p->P_u, *S_u = p->S_u;   /* v tables: leading nv columns, ld nq */
    /* (src/rhs.jl:317-333) */
    cfwd(G1, S_u, P_u, nq, nu, U);
    cfwd(G2, P_u, S_u, nq, nu, U);
    for (int i = 0; i < nq * nq; ++i) { H1[i] = Cw[i] * G1[i]; H2[i] = Cw[i] * G2[i]; }
    cfwd(G1, S_u, P_u, nq, nv, V);
    cfwd(G2, P_u, S_u, nq, nv, V);
    for (int i = 0; i < nq * nq; ++i) { H1[i] += Aw[i] * G1[i]; H2[i] += Aw[i] * G2[i]; }
    cbwd(Rv, S_u, P_u, nq, nv, H1, 0);
    cbwd(Rv, P_u, S_u, nq, nv, H2, 1);
    for (int i = 0; i < nq * nq; ++i) { G1[i] *= Cw[i]; G2[i] *= Cw[i]; }
    cbwd(VB, S_u, P_u, nq, nu, G1, 0);
    cbwd(VB, P_u, S_u, nq, nu, G2, 1);
    /* (src/rhs.jl:340-350) */
    int n2 = nv * nv;
    double om = p->omega;
    double c1 = -(om * om) * ct, c2 = -2.0 * (om * om * om) * st * ct, c3 = 1.0 - ct, c4 = om * st;
    for (int m = 0; m < n2; ++m) {
        long o = m + (long)n2 * e;
        Rv[m] = (-four_hh) * Rv[m] + c1 * p->PiSP[o] + c2 * p->PiBSP2[o] + c3 * p->PiCdiv[o] + c4 * p->PiDdiv[o];
        if (p->src_on) Rv[m] += srct * p->src_projE[o];
    }
    /* (src/rhs.jl:421-473) */
    for (int m = 0; m < nu * nu; ++m) Fu[m] = 0.0;
    const double* Aa[4] = {w->AB, w->AT, w->AL, w->AR};
    const double* Da[4] = {w->DB, w->DT, w->DL, w->DR};
    const double* Fa[4] = {w->FB, w->FT, w->FL, w->FR};
    for (int sd = 0; sd < 4; ++sd) {
        const double* Ev = (sd == 0 || sd == 2) ? p->Ebl_u : p->Ebr_u;
        const double* E1v = (sd == 0 || sd == 2) ? p->E1bl_u : p->E1br_u;
        const double* E1u = E1v;
        double YA[MAXN], YD[MAXN], YF[MAXN];
        long z0 = (long)nq * e;
        for (int b = 0; b < nv; ++b) {
            double sA = 0.0, sD = 0.0;
            for (int i = 0; i < nq; ++i) {
                sA = fma(P_u[i + nq * b], p->w[i] * Aa[sd][z0 + i], sA);
                sD = fma(P_u[i + nq * b], p->w[i] * Da[sd][z0 + i], sD);
            }
            YA[b] = sA; YD[b] = sD;
        }
        for (int b = 0; b < nu; ++b) {
            double sF = 0.0;
            for (int i = 0; i < nq; ++i) sF = fma(P_u[i + nq * b], p->w[i] * Fa[sd][z0 + i], sF);
            YF[b] = sF;
        }
        if (sd < 2) {
            for (int b = 0; b < nv; ++b)
                for (int a = 0; a < nv; ++a) {
                    int k = a + nv * b;
                    Rv[k] = Rv[k] + (two_h * Ev[a]) * YA[b];
                    Rv[k] = Rv[k] + (four_hh * E1v[a]) * YD[b];
                }
            for (int b = 0; b < nu; ++b)
                for (int a = 0; a < nu; ++a) Fu[a + nu * b] = Fu[a + nu * b] + (1.0 * E1u[a]) * YF[b];
        } else {
            for (int b = 0; b < nv; ++b)
                for (int a = 0; a < nv; ++a) {
                    int k = a + nv * b;
                    Rv[k] = Rv[k] + (two_h * YA[a]) * Ev[b];
                    Rv[k] = Rv[k] + (four_hh * YD[a]) * E1v[b];
                }
            for (int b = 0; b < nu; ++b)
                for (int a = 0; a < nu; ++a) Fu[a + nu * b] = Fu[a + nu * b] + (1.0 * YF[a]) * E1u[b];
        }
    }
    /* (src/rhs.jl:361-382) */
    int n = nu * nu;
    double bvec[MAXN * MAXN];
    const double* Ainv = p->Ainv_u + (long)((n * (n + 1)) / 2) * e;
    for (int i = 0; i < n; ++i) { bvec[i] = VB[i] - Fu[i]; UT[i] = 0.0; }
    bvec[0] += V[0];
    for (int jj = 0; jj < n; ++jj) {
        double bj = bvec[jj], s = 0.0;
        int off = (jj * (jj + 1)) / 2;
        for (int i = 0; i < jj; ++i) {
            double a = Ainv[off + i];
            UT[i] = fma(a, bj, UT[i]);
            s = fma(a, bvec[i], s);
        }
        UT[jj] = fma(Ainv[off + jj], bj, UT[jj] + s);
    }
}

static void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

static int rhs_ws(const orc_pre* p, orc_ws* w, const double* u, const double* v, double t,
                  double* ut, double* vt, int cheb_mode, orc_status* st) {
    int N = p->N, nu = p->qu1, nv = p->qv1, qh = p->qh;
    long E = (long)N * N;
    double h = 1.0 / N, two_h = 2.0 / h, four_hh = 4.0 / (h * h);
    double ct = p->mms_on ? cos(p->omega * t) : 0.0;
    double stt = p->mms_on ? sin(p->omega * t) : 0.0;
    double srct = 0.0;
    if (p->src_on && t < p->src_Ts) {
        double sv = sin(M_PI * t / p->src_Ts);
        srct = sin(p->src_omega * t) * (sv * sv);
    }
    memset(st, 0, sizeof(*st));
    /* region 1: traces and Chebyshev intervals */
    long mfail = LONG_MAX;
    int nmax = 1;
#pragma omp parallel for schedule(static) reduction(min : mfail) reduction(max : nmax)
    for (long e = 0; e < E; ++e) {
        double U[MAXN * MAXN], V[MAXN * MAXN];
        load_elem(u, E, nu, e, U);
        load_elem(v, E, nv, e, V);
        trace4(w->duB, w->duT, w->duL, w->duR, p->P_u, nu, nu, p->E1bl_u, p->E1br_u, U, e);
        trace4(w->vB, w->vT, w->vL, w->vR, p->P_u, nu, nv, p->Ebl_u, p->Ebr_u, V, e);
        trace4(w->dvB, w->dvT, w->dvL, w->dvR, p->P_u, nu, nv, p->E1bl_u, p->E1br_u, V, e);
        if (!p->use_cg) {
            int n = cheb_setup_elem(p, V, e, w->G + (long)qh * qh * e, w->th + e, w->de + e, w->s1 + e);
            if (n < 0) { if (e + 1 < mfail) mfail = e + 1; w->ne[e] = 1; }
            else { w->ne[e] = n; if (n > nmax) nmax = n; }
        }
    }
    /* region 2: faces */
    interior_faces(p, w, 1, two_h);
    interior_faces(p, w, 0, two_h);
    long key = 0;
    for (int s = 0; s < 4; ++s) {
        long k2 = boundary_side(p, w, s, ct, stt, two_h);
        if (k2 > 0) key = key == 0 ? k2 : (k2 < key ? k2 : key);
    }
    if (key > 0) {
        st->code = (key % 2 == 1) ? 1 : 2;
        st->key = key;
        st->side = (int)((key - 1) / (2L * nu * N));
        return st->code;
    }
    if (!p->use_cg && mfail != LONG_MAX) { st->code = 3; st->elem = mfail; return 3; }
    st->cheb_steps = nmax;
    /* region 3 */
    long cg_fail = 0, cg_bad = LONG_MAX;
#pragma omp parallel for schedule(static) reduction(+ : cg_fail) reduction(min : cg_bad)
    for (long e = 0; e < E; ++e) {
        double U[MAXN * MAXN], V[MAXN * MAXN], Rv[MAXN * MAXN], UT[MAXN * MAXN], X[MAXN * MAXN];
        load_elem(u, E, nu, e, U);
        load_elem(v, E, nv, e, V);
        element_rhs(p, w, e, U, V, ct, stt, srct, two_h, four_hh, Rv, UT);
        if (p->use_cg) {
            int r = cg_elem(p, V, Rv, X, e);
            if (r < 0) { if (e + 1 < cg_bad) cg_bad = e + 1; }
            else cg_fail += r;
        } else {
            int n = cheb_mode == 1 ? w->ne[e] : nmax;
            cheb_iter_elem(p, w->G + (long)qh * qh * e, w->th[e], w->de[e], w->s1[e], n, Rv, X);
        }
        for (int b = 0; b < nu; ++b)
            for (int a = 0; a < nu; ++a) ut[e + E * (a + (long)nu * b)] = UT[a + nu * b];
        for (int b = 0; b < nv; ++b)
            for (int a = 0; a < nv; ++a) vt[e + E * (a + (long)nv * b)] = X[a + nv * b];
    }
    st->cg_unconverged = cg_fail;
    if (cg_bad != LONG_MAX) {
        st->code = 4;
        st->elem = cg_bad;
        st->j = (cg_bad - 1) % N + 1;
        st->k = (cg_bad - 1) / N + 1;
        return 4;
    }
    return 0;
}

int orc_rhs(const orc_pre* p, const double* u, const double* v, double t, double* ut, double* vt,
            int cheb_mode, int nthreads, orc_status* st) {
    orc_ws w;
    set_threads(nthreads);
    if (ws_alloc(p, &w)) return -1;
    int r = rhs_ws(p, &w, u, v, t, ut, vt, cheb_mode, st);
    ws_free(&w);
    return r;
}

/* (src/solver.jl:289-321) with t accumulated as t += dt (src/solver.jl:213-216) */
int orc_rk4_steps(const orc_pre* p, double* u, double* v, double t0, double dt, long nsteps,
                  int cheb_mode, int nthreads, double* t_end, orc_status* st) {
    long E = (long)p->N * p->N;
    long nU = E * p->qu1 * p->qu1, nV = E * p->qv1 * p->qv1;
    orc_ws w;
    set_threads(nthreads);
    if (ws_alloc(p, &w)) return -1;
    double* buf = (double*)malloc(sizeof(double) * (size_t)(10 * (nU + nV)));
    if (!buf) { ws_free(&w); return -1; }
    double *k1u = buf, *k2u = k1u + nU, *k3u = k2u + nU, *k4u = k3u + nU, *su = k4u + nU;
    double *k1v = su + nU, *k2v = k1v + nV, *k3v = k2v + nV, *k4v = k3v + nV, *sv = k4v + nV;
    double t = t0;
    int r = 0;
    memset(st, 0, sizeof(*st));
    int maxsteps = 0;
    for (long it = 0; it < nsteps; ++it) {
        double hd = 0.5 * dt;
        if ((r = rhs_ws(p, &w, u, v, t, k1u, k1v, cheb_mode, st))) break;
        if (st->cheb_steps > maxsteps) maxsteps = st->cheb_steps;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < nU; ++i) su[i] = u[i] + hd * k1u[i];
#pragma omp parallel for schedule(static)
        for (long i = 0; i < nV; ++i) sv[i] = v[i] + hd * k1v[i];
        if ((r = rhs_ws(p, &w, su, sv, t + 0.5 * dt, k2u, k2v, cheb_mode, st))) break;
        if (st->cheb_steps > maxsteps) maxsteps = st->cheb_steps;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < nU; ++i) su[i] = u[i] + hd * k2u[i];
#pragma omp parallel for schedule(static)
        for (long i = 0; i < nV; ++i) sv[i] = v[i] + hd * k2v[i];
        if ((r = rhs_ws(p, &w, su, sv, t + 0.5 * dt, k3u, k3v, cheb_mode, st))) break;
        if (st->cheb_steps > maxsteps) maxsteps = st->cheb_steps;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < nU; ++i) su[i] = u[i] + dt * k3u[i];
#pragma omp parallel for schedule(static)
        for (long i = 0; i < nV; ++i) sv[i] = v[i] + dt * k3v[i];
        if ((r = rhs_ws(p, &w, su, sv, t + dt, k4u, k4v, cheb_mode, st))) break;
        if (st->cheb_steps > maxsteps) maxsteps = st->cheb_steps;
        double c = dt / 6.0;
        int fin = 1;
#pragma omp parallel for schedule(static) reduction(&& : fin)
        for (long i = 0; i < nU; ++i) {
            u[i] = u[i] + c * (k1u[i] + 2.0 * k2u[i] + 2.0 * k3u[i] + k4u[i]);
            fin = fin && isfinite(u[i]);
        }
#pragma omp parallel for schedule(static) reduction(&& : fin)
        for (long i = 0; i < nV; ++i) {
            v[i] = v[i] + c * (k1v[i] + 2.0 * k2v[i] + 2.0 * k3v[i] + k4v[i]);
            fin = fin && isfinite(v[i]);
        }
        t += dt;
        if (!fin) { st->code = 5; st->step = it + 1; r = 5; break; }
    }
    st->cheb_steps = maxsteps;
    if (t_end) *t_end = t;
    free(buf);
    ws_free(&w);
    return r;
}

int orc_mass_cheb(const orc_pre* p, const double* V, const double* R, double* X, int cheb_mode,
                  int nthreads, orc_status* st) {
    int N = p->N, nv = p->qv1, qh = p->qh;
    long E = (long)N * N;
    set_threads(nthreads);
    memset(st, 0, sizeof(*st));
    double* G = (double*)malloc(sizeof(double) * (size_t)(qh * qh * E + 3 * E));
    int* ne = (int*)malloc(sizeof(int) * (size_t)E);
    double *th = G + qh * qh * E, *de = th + E, *s1 = de + E;
    long mfail = LONG_MAX;
    int nmax = 1;
#pragma omp parallel for schedule(static) reduction(min : mfail) reduction(max : nmax)
    for (long e = 0; e < E; ++e) {
        double Ve[MAXN * MAXN];
        for (int m = 0; m < nv * nv; ++m) Ve[m] = V[m + (long)nv * nv * e];
        int n = cheb_setup_elem(p, Ve, e, G + (long)qh * qh * e, th + e, de + e, s1 + e);
        if (n < 0) { if (e + 1 < mfail) mfail = e + 1; ne[e] = 1; }
        else { ne[e] = n; if (n > nmax) nmax = n; }
    }
    if (mfail != LONG_MAX) { st->code = 3; st->elem = mfail; free(G); free(ne); return 3; }
    st->cheb_steps = nmax;
#pragma omp parallel for schedule(static)
    for (long e = 0; e < E; ++e) {
        int n = cheb_mode == 1 ? ne[e] : nmax;
        cheb_iter_elem(p, G + (long)qh * qh * e, th[e], de[e], s1[e], n, R + (long)nv * nv * e,
                       X + (long)nv * nv * e);
    }
    free(G); free(ne);
    return 0;
}

int orc_mass_cg(const orc_pre* p, const double* V, const double* R, double* X, int nthreads,
                orc_status* st) {
    int N = p->N, nv = p->qv1;
    long E = (long)N * N;
    set_threads(nthreads);
    memset(st, 0, sizeof(*st));
    long bad = LONG_MAX, nf = 0;
#pragma omp parallel for schedule(static) reduction(min : bad) reduction(+ : nf)
    for (long e = 0; e < E; ++e) {
        int r = cg_elem(p, V + (long)nv * nv * e, R + (long)nv * nv * e, X + (long)nv * nv * e, e);
        if (r < 0) { if (e + 1 < bad) bad = e + 1; } else nf += r;
    }
    st->cg_unconverged = nf;
    if (bad != LONG_MAX) {
        st->code = 4; st->elem = bad; st->j = (bad - 1) % N + 1; st->k = (bad - 1) / N + 1;
        return 4;
    }
    return 0;
}

/* Pointwise boundary flux of one side (src/rhs.jl:100-103), exposed for the
 * closed-form checks of test/runtests.jl:295-317. */
void orc_side_flux(const orc_pre* p, int s, double c2f, double af, double cf, double bf, double v_fp,
                   double gun, double gvn, double gu_ex, double gv_ex, double v_ex, double* F1, double* dv) {
    side_flux(p, s, c2f, af, cf, bf, v_fp, gun, gvn, gu_ex, gv_ex, v_ex, F1, dv);
}

int orc_weight_check(double kb, double bf, double v_fp, double v_ex) { return weight_check(kb, bf, v_fp, v_ex); }
