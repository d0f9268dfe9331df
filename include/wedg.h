/* wedg — MI355X (gfx950) right-hand side and RK4 time stepping of the EDG
 * discretization of the Westervelt equation, behind a plain C ABI.
 *
 * The boundary replaces the reference's Julia functions of the hot path
 * (EDG-Westervelt/EDG_Westervelt, WesterveltEDG.jl):
 *   precompute outputs   -> wedg_problem_create      (src/solver.jl:428-694, 716-770: `pre`)
 *   Workspace(pre, N)    -> wedg_workspace_create    (src/rhs.jl:184-220)
 *   edg_rhs!(ws,u,v,t,…) -> wedg_rhs / wedg_rhs_device (src/rhs.jl:490-494)
 *   rk4_step!(u,v,t,dt,…)-> wedg_rk4_step           (src/solver.jl:289-297)
 *   the loop of solve    -> wedg_advance            (src/solver.jl:213-216, 937-940)
 *   _run_source          -> wedg_advance_receivers, wedg_diagnostics, wedg_snapshot
 *                           (src/solver.jl:856-958)
 *   _mms_errors, difference_norms -> wedg_mms_errors, wedg_difference_norms
 *                           (src/solver.jl:772-836)
 *   many solves at once  -> wedg_batch_create, wedg_batch_advance (no counterpart:
 *                           independent instances advanced together)
 * Arrays are the reference's column-major Julia arrays, byte for byte: a Julia
 * caller passes pointer(pre.Cw), pointer(u), … unchanged. The state u is
 * (N, N, q_u+1, q_u+1) and v is (N, N, q_v+1, q_v+1), element e = j + N k
 * fastest. Everything is FP64.
 *
 * Conventions: every function returns a wedg_status; no C++ exception crosses
 * the ABI. One workspace owns one HIP stream and is not re-entrant; different
 * workspaces may be used from different threads. Host-pointer entry points
 * copy in, run and copy out synchronously; *_device entry points take device
 * pointers and are asynchronous on the workspace stream.
 */
#ifndef WEDG_H
#define WEDG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum wedg_status {
    WEDG_OK = 0,
    WEDG_ERR_BOUNDARY_TRACE = 1,  /* 1 - k beta v_h <= 0 at a face point (src/rhs.jl:114) */
    WEDG_ERR_BOUNDARY_DATA = 2,   /* 1 - k beta g <= 0 for the boundary data (src/rhs.jl:115) */
    WEDG_ERR_MASS = 3,            /* mass matrix not positive definite (src/rhs.jl:916) */
    WEDG_ERR_DIMENSION = 4,       /* sizes do not match (DimensionMismatch, src/rhs.jl:500-503) */
    WEDG_ERR_NONFINITE = 5,       /* non-finite state after a step (src/solver.jl:318, 940) */
    WEDG_ERR_HIP = 6,             /* HIP runtime error */
    WEDG_ERR_ARGUMENT = 7,        /* invalid argument or unsupported option */
    WEDG_ERR_SCRATCH = 8,         /* the problem exceeds the static scratch of wedg_rhs_noalloc */
} wedg_status;

/* Boundary kinds, per side (src/solver.jl:37). */
enum { WEDG_ABSORBING_2BETA = 0, WEDG_ABSORBING = 1, WEDG_DIRICHLET = 2, WEDG_WALL = 3 };

/* Parameters of one side (src/solver.jl:39-45). */
typedef struct wedg_side {
    int kind;
    double gamma_abc, eta_abc, sigma, kb, kappa;
} wedg_side;

/* The tables of `precompute` (the fields of the reference's `pre` with the same
 * names and layouts), plus its scalars. Pointers are host memory; they are read
 * during wedg_problem_create only. The v-space tables (P_v, S_v, E*_v) are the
 * leading q_v+1 columns of the u-space ones and are not passed. */
typedef struct wedg_pre {
    int N, qu1, qv1, q_high;
    double omega;                              /* 2 pi sqrt(kx^2 + ky^2) */
    double theta, tau, lambda_u, lambda_v;     /* flux parameters (src/solver.jl:16-21) */
    int mms_on, src_on, one_sided;
    int mass_solver;                           /* 0 = Chebyshev, 1 = conjugate gradients (src/rhs.jl:824-853) */
    double mass_tol;                           /* 1e-14 */
    double src_omega, src_Ts;                  /* GaussianSource omega, Ts (src/solver.jl:80-86) */
    wedg_side bc[4];                           /* bot, top, lft, rgt */
    /* 1D tables */
    const double *w, *P_u, *S_u, *Ebl_u, *Ebr_u, *E1bl_u, *E1br_u, *P_h_v, *W2d_h;
    /* per element */
    const double *Cw, *Aw;                     /* (q_u+1, E, q_u+1) */
    const double *Ainv_u;                      /* (n_u^2 (n_u^2+1)/2, E), packed upper triangle */
    const double *beta_hq;                     /* (q_high, E, q_high) */
    /* per face point: horizontal (q_u+1, N+1, N), vertical (q_u+1, N, N+1) */
    const double *hc2, *haf, *hcf, *hbf, *vc2, *vaf, *vcf, *vbf;
    const double *hc2m, *hafm, *vc2m, *vafm, *hJC, *vJC;
    /* boundary data (q_u+1, N) */
    const double *bcB_sp, *bcB_spy, *bcT_sp, *bcT_spy, *bcL_sp, *bcL_spx, *bcR_sp, *bcR_spx;
    /* projections (q_v+1, q_v+1, N, N) */
    const double *PiSP, *PiBSP2, *PiCdiv, *PiDdiv, *src_projE;
} wedg_pre;

typedef struct wedg_problem wedg_problem;
typedef struct wedg_workspace wedg_workspace;

/* Details of the last failure of a workspace, for the reference's messages. */
typedef struct wedg_error_info {
    int status;         /* wedg_status */
    int side;           /* 0..3 for boundary errors */
    long key;           /* _boundary_key (src/rhs.jl:732) */
    long element;       /* first failing element, linear and 1-based (src/rhs.jl:916) */
    long step;          /* step (1-based, within the last advance) of a failure */
    int cheb_steps;     /* largest Chebyshev step count (conjugate gradients: iterations) of the last call */
    int cg_unconverged; /* conjugate gradients: element systems of the last call left above the
                         * tolerance (the reference warns, src/rhs.jl:855-858) */
} wedg_error_info;

/* Chebyshev step-count policies. */
enum {
    WEDG_CHEB_GLOBAL = 0,   /* every element takes the maximum count over all elements (the reference) */
    WEDG_CHEB_ELEMENT = 1,  /* every element takes its own a-priori count, same 1e-14 bound */
};

/* Uploads the tables of one problem to the current device; replaces `pre`. */
wedg_status wedg_problem_create(const wedg_pre* pre, wedg_problem** out);
wedg_status wedg_problem_destroy(wedg_problem* p);
/* The kernel layout a problem was uploaded with (no counterpart in the
 * reference: it reports what the upload found in precompute's arrays): mode
 * 0 general (a stored u-stiffness inverse per class), 1 separable (c^2
 * constant on each element: the spectral u-solve), 2 uniform (one class, face
 * arrays constant), 3 general with conjugate gradients; the number of
 * coefficient classes; whether the two-lane stage kernels can run it. */
wedg_status wedg_problem_info(const wedg_problem* p, int* mode, int* classes, int* two_lane);

/* Device-resident work arrays and state of one problem (src/rhs.jl:184-220).
 * cheb_policy is WEDG_CHEB_GLOBAL or WEDG_CHEB_ELEMENT. */
wedg_status wedg_workspace_create(wedg_problem* p, int cheb_policy, wedg_workspace** out);
wedg_status wedg_workspace_destroy(wedg_workspace* ws);

/* edg_rhs!: (u_t, v_t) at time t, host arrays (src/rhs.jl:490-550). N, h are
 * checked against the problem (WEDG_ERR_DIMENSION). */
wedg_status wedg_rhs(wedg_workspace* ws, const double* u, const double* v, double t,
                     int N, double h, double* ut, double* vt);

/* The same on device arrays of the current device, asynchronous on the
 * workspace stream; errors of the evaluation are reported by wedg_sync. */
wedg_status wedg_rhs_device(wedg_workspace* ws, const double* u, const double* v, double t,
                            double* ut, double* vt);

/* wedg_rhs_device without a workspace, allocating nothing: no hipMalloc, no
 * hipFree, no allocating library routine. No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The stage
 * kernels' scratch (boundary face slots, failure record, per-wave Chebyshev
 * counts) is one static device array of WEDG_NOALLOC_SCRATCH_BYTES on every
 * GPU the library runs on, held until the process ends; problems with
 * N > WEDG_NOALLOC_MAX_N return WEDG_ERR_SCRATCH. u, v, ut, vt are device
 * arrays of the current device, which must be the problem's. Asynchronous on
 * `stream` (a hipStream_t; NULL is the null stream). When info is not NULL it
 * is DEVICE memory that receives the call's failure record (status WEDG_OK,
 * the largest Chebyshev count) once the stream reaches it. Every call on one
 * GPU shares that GPU's scratch: calls that overlap on one GPU (other streams,
 * other threads) race. */
#define WEDG_NOALLOC_MAX_N 1024
#define WEDG_NOALLOC_SCRATCH_BYTES 950528
wedg_status wedg_rhs_noalloc(const wedg_problem* p, int cheb_policy, const double* u, const double* v,
                             double t, double* ut, double* vt, wedg_error_info* info, void* stream);

/* rk4_step!: one classic RK4 step in place on host arrays (src/solver.jl:289-321).
 * *finite receives all(ws.fin). */
wedg_status wedg_rk4_step(wedg_workspace* ws, double* u, double* v, double t, double dt,
                          int N, double h, int* finite);

/* Device-resident state of the workspace: set, read back, and advance it by
 * nsteps RK4 steps from t0 with t += dt after every step (src/solver.jl:213-216).
 * stop_on_nonfinite = 1 ends the run at the first non-finite step, as the
 * blow-up check of source mode (src/solver.jl:940). */
wedg_status wedg_set_state(wedg_workspace* ws, const double* u, const double* v);
wedg_status wedg_get_state(wedg_workspace* ws, double* u, double* v);
wedg_status wedg_advance(wedg_workspace* ws, double t0, double dt, long nsteps,
                         int stop_on_nonfinite, double* t_end);
/* wedg_advance / wedg_advance_async run whole chunks of chunk_steps steps as
 * one HIP graph: the chunk's stage launches are captured once and replayed
 * with each chunk's arguments (the loop of solve, src/solver.jl:213-216, with
 * RK4 fused into the stage kernels); the results are those of the plain
 * launches, bit for bit. 0 (the default) launches every stage; 1..4096. */
wedg_status wedg_graph_enable(wedg_workspace* ws, long chunk_steps);
/* Lanes per element of the workspace's stage kernels (the Workspace of
 * src/rhs.jl:184 has no counterpart: its element-block split,
 * src/rhs.jl:222-252, is the CPU's form of the same choice). 1 (the default):
 * one lane holds all of an element's arrays; 2: a pair of lanes shares each
 * element, which doubles the waves and halves the registers per lane. The
 * two-lane kernels cover every q_u <= 6 and both pairings (n_v = n_u,
 * n_u - 1) with the spectral u-solve (coefficients constant on each element)
 * and Chebyshev iteration; other problems keep one lane whatever is set
 * (wedg_problem_info says which). Results agree with the
 * one-lane kernels to rounding. A batch runs each workspace with the lanes
 * it had at wedg_batch_create; changing them while the batch exists makes
 * its advance fail with WEDG_ERR_ARGUMENT. wedg_workspace_lanes_per_element
 * reports the lanes the workspace's calls use. */
wedg_status wedg_workspace_set_lanes_per_element(wedg_workspace* ws, int lanes);
wedg_status wedg_workspace_lanes_per_element(const wedg_workspace* ws, int* lanes);
/* The same without waiting: the caller calls wedg_sync before reading results. */
wedg_status wedg_advance_async(wedg_workspace* ws, double t0, double dt, long nsteps,
                               int stop_on_nonfinite, double* t_end);
wedg_status wedg_sync(wedg_workspace* ws);

/* Device pointers of the workspace state (u, v) for callers that own the data flow. */
wedg_status wedg_state_device(wedg_workspace* ws, double** u, double** v);

wedg_status wedg_last_error(const wedg_workspace* ws, wedg_error_info* info);

/* Source mode (src/solver.jl:872-958). */

/* The coefficients of the energy diagnostics: geo.err_c2 and geo.err_beta,
 * (N, N, q_u+1, q_u+1) as built by the reference's geometry (src/solver.jl:224),
 * and the 2D weights geo.W2d ((q_u+1)^2). Called once per problem before
 * wedg_diagnostics. */
wedg_status wedg_problem_set_diag(wedg_problem* p, const double* err_c2, const double* err_beta,
                                  const double* W2d);

/* wedg_advance with the receiver time series of _run_source (src/solver.jl:899-905,
 * 938-941): after every step, v_h at each receiver, eta_r^T V[j_r, k_r] xi_r, into
 * recv[step * nrecv + r] (nsteps x nrecv). recv_j/recv_k are the 0-based element
 * row/column of locate(y_r)/locate(x_r) and rows holds per receiver the modal rows
 * vrow(eta_r) then vrow(xi_r), 2 (q_v+1) values. */
wedg_status wedg_advance_receivers(wedg_workspace* ws, double t0, double dt, long nsteps, int nrecv,
                                   const int* recv_j, const int* recv_k, const double* rows,
                                   int stop_on_nonfinite, double* recv, double* t_end);

/* The record-step diagnostics of the current state (src/solver.jl:856-889, 893-896,
 * 907-918): per element (e = j + N k) the energy (h^2/4) sum W2d (1/2 c^2 |grad u_h|^2
 * + (1/2 - 2/3 beta v_h) v_h^2), max |v_h| and min (1 - 2 beta v_h) at the points,
 * into E, vmax, ndmin (N^2 each, for the caller's serial accumulation in row-j,
 * column-k order), and the slice P_v (eta0^T V[j0, k])^T for k = 0..N-1 into slice
 * (N (q_u+1) values). eta0 is vrow(eta) of locate(slice_y), q_v+1 values. */
wedg_status wedg_diagnostics(wedg_workspace* ws, double* E, double* vmax, double* ndmin, int j0,
                             const double* eta0, double* slice);

/* The snapshot vgrid of _run_source (src/solver.jl:919-923): P_v V[j, k] P_v^T of
 * every element into grid, (N (q_u+1)) x (N (q_u+1)) column-major, row j (q_u+1) + p
 * (the y points), column k (q_u+1) + q (the x points). */
wedg_status wedg_snapshot(wedg_workspace* ws, double* grid);

/* Error norms (src/solver.jl:772-836), after wedg_problem_set_diag. Both give per
 * element e = j + N k the quadrature sums elem[c * N^2 + e] (c = 0..3) that the
 * reference accumulates over j outer, k inner before h/2 sqrt(.):
 *
 * _mms_errors (src/solver.jl:774-810) of the workspace state: W (u* - u_h)^2,
 * W (v* - v_h)^2, W |grad u* - grad_h u_h|^2 and W (1/2 c^2 |grad u* - grad_h u_h|^2
 * + 1/2 (v* - v_h)^2) with u* = cu Phi, v* = cv Phi, cu = cos(wT) - 1,
 * cv = -w sin(wT), Phi = cos(tpx x + px) sin(tpy y + py), tpx = 2 pi kx,
 * tpy = 2 pi ky. cx, sx hold cos and sin of tpx x + px at the x points of element
 * column k ([k][q], N (q_u+1) values), cy, sy those of tpy y + py at the y points
 * of element row j ([j][p]).
 *
 * difference_norms (src/solver.jl:819-836) of two host states A, B (the
 * workspace's layout): the sums of e = A - B (the third is the gradient sum,
 * which the reference does not report). */
wedg_status wedg_mms_errors(wedg_workspace* ws, double cu, double cv, double tpx, double tpy,
                            const double* cx, const double* sx, const double* cy, const double* sy,
                            double* elem);
wedg_status wedg_difference_norms(wedg_workspace* ws, const double* uA, const double* vA,
                                  const double* uB, const double* vB, double* elem);

/* Kernel timing with HIP events on the workspace stream over the next run(s):
 * enable, then read the accumulated per-kernel-family milliseconds and counts.
 * family: 0 = k_prep (Chebyshev count and boundary slots of the first stage of a call),
 * 1 = k_stage (one RK4 stage: traces, faces, volume terms, source, lifts, u-solve,
 * mass solve, the RK4 update and the next stage's count and boundary slots);
 * 2 and 3 are not used. */
wedg_status wedg_timing_enable(wedg_workspace* ws, int on);
wedg_status wedg_timing_read(wedg_workspace* ws, int family, double* ms, long* launches);

/* Chebyshev steps of the last synchronized call: the mean over stages and
 * waves of the count each wave of 64 elements ran, and the number of stages. */
wedg_status wedg_cheb_stats(const wedg_workspace* ws, double* mean_steps, long* stages);

/* Measurement utility (no counterpart in the reference): a plain copy kernel of
 * n doubles run `repeats` times on the current device, for calibrating the
 * byte counters of a profiler; *ms receives the average time per copy. */
wedg_status wedg_probe_copy(size_t n, int repeats, double* ms);

/* Batched instances. The reference advances one instance per call (the loop
 * of solve, src/solver.jl:213-216); a batch advances n independent workspaces
 * (uneven N, q_u, q_v, coefficients, boundary kinds, time steps and step
 * counts) together: one kernel launch per RK stage for each group of
 * instances with the same (q_u, q_v), coefficient layout and lanes per
 * element (wedg_workspace_set_lanes_per_element), on the batch's streams. Instance i runs nsteps[i] steps of dt[i] from t0[i] exactly as
 * wedg_advance(ws[i], t0[i], dt[i], nsteps[i], ...) would (same stage
 * arguments, bitwise the same state); status[i] and t_end[i] (either may be
 * null) receive its status and final time, wedg_last_error(ws[i]) its failure
 * details; the return value is the first non-zero status. A workspace may
 * appear in a batch once; it must not be used elsewhere while the batch runs,
 * and it must outlive the batch. A batch, like a workspace, is not re-entrant.
 * wedg_batch_time gives the device time of the last advance in ms. */
typedef struct wedg_batch wedg_batch;
wedg_status wedg_batch_create(wedg_workspace* const* ws, int n, wedg_batch** out);
wedg_status wedg_batch_destroy(wedg_batch* b);
wedg_status wedg_batch_advance(wedg_batch* b, const double* t0, const double* dt, const long* nsteps,
                               int stop_on_nonfinite, double* t_end, int* status);
wedg_status wedg_batch_time(const wedg_batch* b, double* ms);
/* Source mode in a batch: wedg_set_receivers stores a workspace's receivers
 * (the arguments of wedg_advance_receivers); wedg_batch_advance_receivers is
 * wedg_batch_advance that also records, for every instance i with receivers
 * and recv[i] non-null, v_h at its receivers after each of its nsteps[i]
 * steps into recv[i] ([step][receiver]), as wedg_advance_receivers does for
 * one workspace (src/solver.jl:898-905, 938-941). */
wedg_status wedg_set_receivers(wedg_workspace* ws, int nrecv, const int* recv_j, const int* recv_k,
                               const double* rows);
wedg_status wedg_batch_advance_receivers(wedg_batch* b, const double* t0, const double* dt, const long* nsteps,
                                         int stop_on_nonfinite, double* t_end, int* status, double* const* recv);

/* Launch parameters of a batch (they change no result): the number of streams
 * the groups are dealt to (1..16; default: the process's hardware queues,
 * GPU_MAX_HW_QUEUES or 4), the steps whose stage arguments are staged per
 * copy (0: about 4 MB of arguments), and the schedule: 0 (default) starts
 * every instance at the advance's first step, 1 ends them all at its last
 * (instance i runs its nsteps[i] steps as the last ones of the advance).
 * wedg_batch_config reads them back. */
wedg_status wedg_batch_configure(wedg_batch* b, int lanes, long chunk_steps, int schedule);
wedg_status wedg_batch_config(const wedg_batch* b, int* lanes, long* chunk_steps, int* schedule);

/* ------------------------------------------------------------------------
 * Discrete adjoint of source-mode runs (no counterpart in the reference: the
 * reference's hot path is the forward map edg_rhs! + rk4_step! in
 * mode = :source, src/rhs.jl:490-550, src/solver.jl:289-321, 892-958, whose
 * derivatives these entry points compute).
 *
 * The problem: mode = :source with coefficients constant on every element,
 * c^2, alpha and beta given per element as (N, N) arrays (row j in y, column k
 * in x, element e = j + N k), zero initial state, t += dt after every step
 * (src/solver.jl:213-216, 937-939). The objective is the misfit of u_h on the
 * bottom side y = 0 at its face quadrature points after every RK4 step,
 *   J = 1/2 sum_{n=1..S} sum_k sum_i (h/2) w_i (u_h(x_ki, 0, t_n) - g_n[k, i])^2,
 * with the target g ([n][k][i], q_u+1 points per element) set by
 * wedg_adj_set_target. wedg_adj_gradient returns J and dJ/dc^2, dJ/dbeta per
 * element by the discrete adjoint of RK4 and of every stage (faces, boundary
 * fluxes with their sqrt(1 - k beta v) weights, volume terms, lifts, the
 * u-solve and the mass solve, the latter as the exact solve M^-1 with the
 * Chebyshev iteration to 1e-14), with binomial checkpointing of the RK4
 * states in `snapshots` state slots (the initial state counts as one). */
typedef struct wedg_adj_desc {
    int N, qu1, qv1;
    double theta, tau, lambda_u, lambda_v;     /* flux parameters (src/solver.jl:16-21) */
    wedg_side bc[4];                           /* bot, top, lft, rgt (src/solver.jl:39-45) */
    double mass_tol;                           /* 1e-14 */
    double src_omega, src_Ts;                  /* GaussianSource omega, Ts (src/solver.jl:80-86) */
    /* the 1D tables of wedg_pre */
    const double *w, *P_u, *S_u, *Ebl_u, *Ebr_u, *E1bl_u, *E1br_u, *P_h_v, *W2d_h;
    const double* src_projE;                   /* (q_v+1, q_v+1, N, N): P_h^T (W_h o g) P_h (src/solver.jl:656-660) */
    const double* alpha;                       /* (N, N) per element */
} wedg_adj_desc;

typedef struct wedg_adj wedg_adj;

/* Counts of the last wedg_adj_gradient: forward steps advanced (first sweep and
 * recomputations), steps taped and reversed, snapshots written and the most
 * slots held at once. */
typedef struct wedg_adj_stats {
    long forward_steps, taped_steps, adjoint_steps, snapshots_taken;
    int snapshots_used, tapes;     /* tapes: steps the forward stream may run ahead of the adjoint */
    double enqueue_ms, total_ms;   /* host time to issue the run, and to its end */
} wedg_adj_stats;

/* cheb_policy of the forward steps as in wedg_workspace_create. The
 * coefficients start as c^2 = 1, beta = 0 until wedg_adj_set_params. */
wedg_status wedg_adj_create(const wedg_adj_desc* d, int cheb_policy, wedg_adj** out);
wedg_status wedg_adj_destroy(wedg_adj* a);
/* c2 (> 0) and beta per element, (N, N) each; rebuilds the forward tables. */
wedg_status wedg_adj_set_params(wedg_adj* a, const double* c2, const double* beta);
/* The target g of the misfit for nsteps steps, [n][k][i] (nsteps x N x (q_u+1)). */
wedg_status wedg_adj_set_target(wedg_adj* a, long nsteps, const double* target);
/* nsteps RK4 steps of dt from the zero state: J (needs a target of at least
 * nsteps rows; null: not formed) and the bottom traces u_h after every step
 * into traces ([n][k][i], or null). */
wedg_status wedg_adj_forward(wedg_adj* a, double dt, long nsteps, double* J, double* traces);
/* J and its gradient, dJ_dc2 and dJ_dbeta (N, N) each, by binomial
 * checkpointing with `snapshots` state slots. Three streams (forward, tape,
 * adjoint) run concurrently; the forward stream runs ahead of the adjoint by
 * as many reversed steps as there are tapes (as many as fit in 2 GiB, 4 to 32,
 * allocated on first use). The streams should have hardware queues of their
 * own (GPU_MAX_HW_QUEUES >= 4 plus the process's other live streams). */
wedg_status wedg_adj_gradient(wedg_adj* a, double dt, long nsteps, int snapshots, double* J, double* dJ_dc2,
                              double* dJ_dbeta);
wedg_status wedg_adj_last_stats(const wedg_adj* a, wedg_adj_stats* st);
/* Lanes per element of the instance's forward stage kernels (its forward
 * runs, checkpoint recomputations and taped steps alike; the adjoint kernels
 * keep one lane), as wedg_workspace_set_lanes_per_element: 1 (the default) or
 * 2 where the two-lane kernels admit the problem. No counterpart in the
 * reference, whose gradient follows its CPU forward. Results agree with one
 * lane to rounding; recomputed states stay bitwise the forward's. */
wedg_status wedg_adj_set_lanes_per_element(wedg_adj* a, int lanes);
wedg_status wedg_adj_lanes_per_element(const wedg_adj* a, int* lanes);
/* One right-hand side (u_t, v_t) = F(u, v, t; c^2, beta) differentiated: the
 * vector-Jacobian product, from the cotangents (ub, vb) of (u_t, v_t) those of
 * (u, v, c^2, beta); and the tangent-linear map, (du_t, dv_t) along
 * (du, dv, dc2, dbeta). Host arrays in the state layouts. */
wedg_status wedg_adj_rhs_vjp(wedg_adj* a, const double* u, const double* v, double t, const double* ub,
                             const double* vb, double* u_bar, double* v_bar, double* c2_bar, double* beta_bar);
wedg_status wedg_adj_rhs_tl(wedg_adj* a, const double* u, const double* v, double t, const double* du,
                            const double* dv, const double* dc2, const double* dbeta, double* dut, double* dvt);

/* Several instances' forward runs and gradients together (no counterpart in
 * the reference): each instance runs what wedg_adj_forward / wedg_adj_gradient
 * run on it, with identical results, in lockstep rounds on one stream; per
 * round, the instances with the same (q_u, q_v, kernel layout, lanes per
 * element) share one launch for every RK stage and every adjoint kernel. Instances keep their
 * coefficients and targets (wedg_adj_set_params, wedg_adj_set_target) and may
 * not be used elsewhere while a batch call runs. status[i] (or null) receives
 * instance i's status; the return value is the first failure. Per-instance
 * counts: wedg_adj_last_stats. */
typedef struct wedg_adj_batch wedg_adj_batch;
wedg_status wedg_adj_batch_create(wedg_adj* const* a, int n, wedg_adj_batch** out);
wedg_status wedg_adj_batch_destroy(wedg_adj_batch* b);
/* J[i] after nsteps[i] steps of dt[i] from the zero state (a target of at least nsteps[i] rows each). */
wedg_status wedg_adj_batch_forward(wedg_adj_batch* b, const double* dt, const long* nsteps, double* J, int* status);
/* J[i] and dJ_dc2[i], dJ_dbeta[i] ((N_i, N_i) host arrays) with `snapshots` slots per instance.
 * Each kernel group runs its forward and its adjoint launches on two streams
 * of its own, the adjoint one round behind (two tapes per instance). */
wedg_status wedg_adj_batch_gradient(wedg_adj_batch* b, const double* dt, const long* nsteps, int snapshots, double* J,
                                    double* const* dJ_dc2, double* const* dJ_dbeta, int* status);
/* Device time of the last batch call (stream events) and the host time to issue it. */
wedg_status wedg_adj_batch_time(const wedg_adj_batch* b, double* device_ms, double* enqueue_ms);

/* Library and device facts: compiled architecture, device name, CU count. */
const char* wedg_version(void);
wedg_status wedg_device_info(char* name, size_t len, int* cu_count, double* hbm_gbs);

#ifdef __cplusplus
}
#endif
#endif
