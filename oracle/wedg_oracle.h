/* CPU oracle for the EDG Westervelt right-hand side and RK4 (test infrastructure).
 *
 * This header belongs to the oracle under oracle/: a plain-C, FP64 restatement
 * of the reference's batched algorithm (src/rhs.jl, src/solver.jl:278-321) used
 * only by tests/, the smoke check and the CPU-baseline leg of bench.py, as the
 * checker. The product path never links or calls it.
 *
 * All arrays use the reference's column-major layouts (the tables of
 * `precompute`, src/solver.jl:428-770); the state u is (N, N, n_u, n_u) with
 * element e = j + N k fastest.
 */
#ifndef WEDG_ORACLE_H
#define WEDG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_pre {
    int N, qu1, qv1, qh;
    double omega, theta, tau, lambda_u, lambda_v;
    int mms_on, src_on, one_sided, use_cg;
    double mass_tol;
    double src_omega, src_Ts;          /* source time factor sin(w t) sin^2(pi t / Ts), t < Ts */
    int kind[4];                       /* 0 absorbing_2beta, 1 absorbing, 2 dirichlet, 3 wall (bot, top, lft, rgt) */
    double gam[4], eta[4], sig[4], kb[4], kap[4];
    const double *w, *P_u, *S_u, *Ebl_u, *Ebr_u, *E1bl_u, *E1br_u, *P_h_v, *W2d_h;
    const double *Cw, *Aw, *Ainv_u, *beta_hq, *beta_hE;
    const double *hc2, *haf, *hcf, *hbf, *vc2, *vaf, *vcf, *vbf;
    const double *hc2m, *hafm, *vc2m, *vafm, *hJC, *vJC;
    const double *bcB_sp, *bcB_spy, *bcT_sp, *bcT_spy, *bcL_sp, *bcL_spx, *bcR_sp, *bcR_spx;
    const double *PiSP, *PiBSP2, *PiCdiv, *PiDdiv, *src_projE;
} orc_pre;

/* Outcome of one evaluation: code 0 ok; 1 boundary weight of the trace, 2 of the
 * data (side = 0..3, key = the reference's _boundary_key); 3 Chebyshev mass
 * matrix (elem = first failing linear element, 1-based); 4 CG mass matrix
 * (j, k 1-based); 5 non-finite state after an RK4 step (step = 1-based). */
typedef struct orc_status {
    int code, side;
    long key;
    long elem, j, k, step;
    int cheb_steps;                    /* Chebyshev steps taken (max over elements) */
    long cg_unconverged;
} orc_status;

/* cheb_mode: 0 = one step count for all elements, the maximum (the reference);
 *            1 = each element takes its own a-priori count. */
int orc_rhs(const orc_pre* p, const double* u, const double* v, double t,
            double* ut, double* vt, int cheb_mode, int nthreads, orc_status* st);

int orc_rk4_steps(const orc_pre* p, double* u, double* v, double t0, double dt, long nsteps,
                  int cheb_mode, int nthreads, double* t_end, orc_status* st);

/* The mass systems alone (src/rhs.jl:862-868, 942-951): X = M(V)^-1 R. */
int orc_mass_cheb(const orc_pre* p, const double* V, const double* R, double* X,
                  int cheb_mode, int nthreads, orc_status* st);
int orc_mass_cg(const orc_pre* p, const double* V, const double* R, double* X,
                int nthreads, orc_status* st);

/* Pointwise boundary flux of side s (0..3) with the parameters in p->kind/gam/… */
void orc_side_flux(const orc_pre* p, int s, double c2f, double af, double cf, double bf, double v_fp,
                   double gun, double gvn, double gu_ex, double gv_ex, double v_ex, double* F1, double* dv);
int orc_weight_check(double kb, double bf, double v_fp, double v_ex);

#ifdef __cplusplus
}
#endif
#endif
