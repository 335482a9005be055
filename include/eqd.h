/* libeqd -- MI355X (gfx950) split-node fault-stiffness engine and BP8 RHS, C ABI.
 *
 * The library replaces the hot path of EarthquakeDiffinitive.jl: the split-node CG
 * solver (ElasticitySplitNode.CGSolver) and the dense fault stiffness build
 * (FaultResponse.fault_stiffness, :exact and :toeplitz), plus the BP8 right-hand
 * side. All signatures use plain pointers and sizes so a Julia host can `ccall`
 * them; host arrays are Julia's column-major Float64 layout (pointer(K) can be
 * passed directly).
 *
 * Conventions
 *   - Every function returns an eqd_status; 0 is success. On failure
 *     eqd_last_error() holds a message (thread-local).
 *   - Full-space DOF vectors have the reference's layout: [u-(3N); u+(3N)],
 *     component-major, x1 fastest (src/ElasticitySplitNode.jl:36-42).
 *   - Slip/traction vectors are indexed over Omega_f nodes, x2 fastest
 *     (src/FaultResponse.jl:112-126). K is 2nf x 2nf, column-major, rows
 *     [dtau2; dtau3], columns [s2; s3] (src/FaultResponse.jl:214-232).
 *   - A solver handle is single-threaded (like Krylov's CgWorkspace); distinct
 *     handles may be used from different threads. Handles own their device memory.
 */
#ifndef EQD_H
#define EQD_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef int eqd_status;
enum {
    EQD_OK = 0,
    EQD_ERR_ARG = 1,        /* invalid argument (ErrorException in the reference) */
    EQD_ERR_GRID = 2,       /* grid/patch legality (src/FaultResponse.jl:118-124, src/BP8.jl:189-204) */
    EQD_ERR_DEVICE = 3,     /* HIP runtime failure, or no usable gfx950 device */
    EQD_ERR_NOMEM = 4,      /* device allocation failed */
    EQD_ERR_RANGE = 5,      /* rhs not in range(P) */
    EQD_ERR_NEWTON = 6      /* rate-state Newton solve did not converge (src/RateStateFriction.jl:93-94) */
};

typedef struct eqd_op eqd_op;
typedef struct eqd_solver eqd_solver;

/* Two SBP blocks [-L_normal,0]x[-L_fault,L_fault]^2 and [0,L_normal]x[...]^2 with
 * n1 x n23 x n23 nodes each (src/BP8.jl:196-206); Omega_f = (-l_f, l_f)^2. */
typedef struct {
    int n1, n23;
    double L_normal, L_fault;
    double lambda, mu;
    double l_f;
} eqd_grid_desc;

typedef struct {
    int64_t nf, n2f, n3f;   /* Omega_f nodes and its two extents */
    int64_t ntot;           /* full-space DOFs, 6*n1*n23^2 */
    int64_t nreduced;       /* unknowns the device stores per column (range(P), interior) */
    double h1, h2;
} eqd_op_info;

enum { EQD_PREC_FP64 = 0, EQD_PREC_MIXED = 1 };
enum { EQD_PRECOND_NONE = 0, EQD_PRECOND_JACOBI = 1 };
enum { EQD_ENGINE_BATCHED = 0, EQD_ENGINE_PLAIN = 1 };

typedef struct {
    double rtol;              /* default 1e-10 (src/ElasticitySplitNode.jl:365) */
    double atol;              /* default 0 */
    int64_t itmax;            /* 0 = 2*ntot, Krylov's default */
    int precond;              /* EQD_PRECOND_NONE or EQD_PRECOND_JACOBI (M^{-1} = diag(A)^{-1}, src/ElasticitySplitNode.jl:343-366) */
    int precision;            /* EQD_PREC_FP64 or EQD_PREC_MIXED (FP32 CG inside FP64 reliable updates) */
    int engine;               /* EQD_ENGINE_BATCHED (device-side alpha/beta) or _PLAIN (host alpha/beta, one column) */
    int tile_k;               /* columns interleaved per tile: 1, 8, 16, 32 or 64; 0 = autotuned: the first
                                 build times 16 CG iterations of its own first batch at each width */
    int max_cols_in_flight;   /* columns resident at once (0 = all, bounded by memory) */
    int reliable_interval;    /* mixed precision: iterations between FP64 residual refreshes (0 = 48) */
    int graph_iters;          /* iterations captured per HIP graph replay (0 = 8) */
    int timing;               /* 1: no graphs, HIP events around every kernel (measurement runs) */
} eqd_cg_opts;

typedef struct {
    int64_t solves, iterations, max_iterations, unconverged;
    double mean_iterations;
    int precond;
    int tile_k;               /* tile width of the last build (the autotuned one when tile_k was 0) */
} eqd_report;

void eqd_cg_opts_default(eqd_cg_opts* o);

/* Operator: assembles the 1D tables of -HP(D+SAT)P, HP(D+SAT) and the traction map.
 * Replaces split_node_system + FaultElasticity setup (src/ElasticitySplitNode.jl:106-214,
 * src/FaultResponse.jl:96-146). Errors: n1 or n23 < 9, no grid node on +-l_f. */
eqd_status eqd_operator_create(const eqd_grid_desc* g, eqd_op** out);
eqd_status eqd_operator_destroy(eqd_op* op);
eqd_status eqd_operator_info(const eqd_op* op, eqd_op_info* info);
/* Omega_f coordinates (x2f[n2f], x3f[n3f]); fault_grid_axes (src/FaultResponse.jl:161). */
eqd_status eqd_fault_grid_axes(const eqd_op* op, double* x2f, double* x3f);

/* y = A x with A = -HP(D+SAT)P, full-space FP64 vectors (A's product as used by cg!). */
eqd_status eqd_apply_A(eqd_op* op, const double* x, double* y);
/* rhs = HP(D+SAT) chi(s) for slip on Omega_f (src/FaultResponse.jl:202-203), full space. */
eqd_status eqd_build_rhs(eqd_op* op, const double* s2, const double* s3, double* rhs);

/* jacobi_preconditioner(A) (src/ElasticitySplitNode.jl:343-357): minv[n] = diag(A)^{-1} of
 * A = -HP(D+SAT)P over the full space (n = 6 n1 n23^2), A's zero diagonal entries (the
 * far-field rows) floored to 1. Host only (from the operator tables; no device needed). */
eqd_status eqd_jacobi_preconditioner(const eqd_grid_desc* g, double* minv, int64_t n);

/* CGSolver(A; rtol, atol, itmax, precond) (src/ElasticitySplitNode.jl:365). precond JACOBI
 * runs preconditioned CG with Krylov's cg! semantics under M: gamma = r.M^{-1}r, stop at
 * sqrt(gamma) <= atol + rtol sqrt(b.M^{-1}b). */
eqd_status eqd_solver_create(eqd_op* op, const eqd_cg_opts* opts, eqd_solver** out);
eqd_status eqd_solver_destroy(eqd_solver* s);
/* duplicate(solver): shares the operator, fresh workspace and counters (:408). */
eqd_status eqd_solver_duplicate(const eqd_solver* s, eqd_solver** out);
/* merge_stats!(into, from) (:426). */
eqd_status eqd_solver_merge_stats(eqd_solver* into, const eqd_solver* from);
/* solver_report(solver) (:442). */
eqd_status eqd_solver_report(const eqd_solver* s, eqd_report* r);

/* split_node_solve(solver, rhs) -> u (:378). rhs must lie in range(P), as every rhs
 * HP(D+SAT)chi does; otherwise EQD_ERR_RANGE. u is the full-space iterate (P u = u). */
eqd_status eqd_split_node_solve(eqd_solver* s, const double* rhs, double* u);

/* shear_traction(fe, s2, s3) -> (dtau2, dtau3) (src/FaultResponse.jl:194-211). */
eqd_status eqd_shear_traction(eqd_solver* s, const double* s2, const double* s3, double* dtau2, double* dtau3);

/* No counterpart in the reference: eqd_shear_traction for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The same one-column solve and the same kernels (results equal to eqd_shear_traction's on a
 * solver with the same options), but the Krylov workspace lives in a static 64 MiB device array
 * (one per device, held until the process ends) and the iterations are launched directly: no
 * hipMalloc, hipFree, graph, event or pinned host memory. Batched engine only. A grid whose
 * one-column workspace exceeds 64 MiB (about 1.6M stored unknowns in FP64, 2M in mixed
 * precision) is refused with EQD_ERR_RANGE. Calls that overlap on one device share the array
 * and race. */
eqd_status eqd_shear_traction_noalloc(eqd_solver* s, const double* s2, const double* s3, double* dtau2,
                                      double* dtau3);

enum { EQD_STIFFNESS_EXACT = 0, EQD_STIFFNESS_TOEPLITZ = 1 };
/* fault_stiffness(fe) / fault_stiffness_toeplitz(fe) (src/FaultResponse.jl:229-287, 339-393).
 * EXACT: cols == NULL builds all 2nf columns into K (ldK >= 2nf, 2nf columns);
 *        otherwise only columns cols[0..ncols) (0-based) are built, into K[:, 0..ncols).
 * TOEPLITZ: 10 solves, full K (cols must be NULL). */
eqd_status eqd_fault_stiffness(eqd_solver* s, int mode, const int* cols, int ncols, double* K, int64_t ldK);
/* Same, with K left in device memory (caller-owned device pointer). */
eqd_status eqd_fault_stiffness_device(eqd_solver* s, int mode, const int* cols, int ncols, double* K_dev,
                                      int64_t ldK);

/* Timing hooks for measurement: device time (ms) of the last stiffness build's
 * iteration phase, and the number of stencil launches and tile-iterations in it. */
eqd_status eqd_solver_last_timing(const eqd_solver* s, double* ms, int64_t* stencil_launches,
                                  int64_t* tile_iterations);

/* Measurement hook, no counterpart in the reference: per-kernel device time of the last
 * stiffness build run with opts.timing = 1. kernel: 0 stencil q = A p, 1 x/r update,
 * 2 p update, 3 per-column finalize, 4 reliable-update accumulate, 5 reliable-update
 * FP64 residual. Returns summed ms, launch count and the sum over launches of the
 * number of tiles with an active column (each tile holds tile_k columns); for kernel 5, of
 * the tiles whose FP64 residual the pass recomputed. */
eqd_status eqd_solver_kernel_timing(const eqd_solver* s, int kernel, double* total_ms, int64_t* launches,
                                    int64_t* active_tiles);

/* Diagnostics, no counterpart in the reference: the host-built 1D tables behind the
 * kernels, available without a device (for inspection and CPU-side tests of the
 * table construction). which: 0 AXs, 1 AXo, 2 MX, 3 Ts, 4 To, 5 D1, 6 D1s,
 * 7 D1full, 8 D1ext (banded: start[nrows], width[nrows], coef[nrows*16]);
 * 100 Hx, 101 Hy, 102 Jt, 103 Jn, 104 Dn (vectors into coef), 105 Jext (scalar).
 * Any output pointer may be NULL. Returns the row count / vector length, or -1. */
int64_t eqd_tables_export(const eqd_grid_desc* g, int which, int* start, int* width, double* coef);

/* ------------------------------------------------------------------------------
 * BP8 time loop on the device: K resident in HBM, one kernel per Tsit5 stage
 * (K row GEMV + rate-state Newton + pore-pressure operator + stage combination),
 * the step controller and the saveat interpolation on the device.
 * ------------------------------------------------------------------------------ */
typedef struct eqd_bp8 eqd_bp8;

/* BP8Params (Table 1, src/BP8.jl:50-87); lambda = 2 mu nu / (1 - 2 nu), eta = mu / (2 c_s),
 * q0 = Q0 / L_fwid are derived as build_model does. */
typedef struct {
    double mu, nu, rho, c_s;
    double a, b, D_RS, V_star, f_star;
    double sigma0, tau_init, V_init, V_zero;
    double Q0, L_gauss, L_fwid, beta, phi, k, viscosity, alpha, S_well, r_well;
    double l_f, dz, t_off, t_f, sigma_min;
} eqd_bp8_params;

enum { EQD_INJECTION_GAUSSIAN = 0, EQD_INJECTION_PEACEMAN = 1 };

void eqd_bp8_default_params(eqd_bp8_params* p);

/* build_model(...) minus the K build (src/BP8.jl:184-258): A_p, source, weights, active
 * ring, tau0, Peaceman WI and cell volume around a given K (2nf x 2nf, column-major,
 * leading dimension ldK; host memory for eqd_bp8_create, device memory for
 * eqd_bp8_create_device). n2f = n3f = sqrt(nf). Errors as build_model: n2f < 9,
 * unknown injection, r_e <= r_well. */
eqd_status eqd_bp8_create(const eqd_bp8_params* p, const double* K, int64_t ldK, int n2f, int injection,
                          eqd_bp8** out);
eqd_status eqd_bp8_create_device(const eqd_bp8_params* p, const double* K_dev, int64_t ldK, int n2f,
                                 int injection, eqd_bp8** out);
/* build_model with the K build on the device (stiffness = EXACT or TOEPLITZ through the
 * solver's engine); K never leaves device memory. */
eqd_status eqd_bp8_create_from_solver(eqd_solver* s, int stiffness, const eqd_bp8_params* p, int injection,
                                      eqd_bp8** out);
eqd_status eqd_bp8_destroy(eqd_bp8* m);
/* nf and the state length 4nf (+1 with the Peaceman well) (src/BP8.jl:262,362-364). */
eqd_status eqd_bp8_sizes(const eqd_bp8* m, int* nf, int* state_len);
/* The model's K (2nf x 2nf column-major, ldK >= 2nf) copied to the host. */
eqd_status eqd_bp8_stiffness(const eqd_bp8* m, double* K, int64_t ldK);
/* A_p as a dense nf x nf row-major matrix, the source and the quadrature weights (host). */
eqd_status eqd_bp8_operators(const eqd_bp8* m, double* Ap, double* source, double* weights);
/* initial_state(m) (src/BP8.jl:337-345): [0; 0; ln theta0; 0 (; 0)]. */
eqd_status eqd_bp8_initial_state(const eqd_bp8* m, double* u);
/* evaluate!(m, u, t) (src/BP8.jl:355-396): V2, V3, |V|, tau2, tau3 on Omega_f (any may be
 * NULL). Like the reference it updates the model's Newton warm start and stress floor
 * statistics. EQD_ERR_NEWTON if a node's Newton solve does not converge. */
eqd_status eqd_bp8_evaluate(eqd_bp8* m, const double* u, double t, double* V2, double* V3, double* Vmag,
                            double* tau2, double* tau3);
/* rhs!(du, u, m, t) (src/BP8.jl:403-433). */
eqd_status eqd_bp8_rhs(eqd_bp8* m, const double* u, double t, double* du);
/* effective_stress_report(m) (src/BP8.jl:287): the smallest sigma0 - p seen before the
 * floor, the number of floor hits (evaluations x nodes) and the mask of nodes ever floored
 * (floor_mask[nf], may be NULL). */
eqd_status eqd_bp8_stress_report(const eqd_bp8* m, double* sigma_lowest, int64_t* floor_hits,
                                 unsigned char* floor_mask);
/* Resets the evaluate! cache (warm start 1e-12, statistics), like a fresh Cache(nf). */
eqd_status eqd_bp8_reset_cache(eqd_bp8* m);

typedef struct {
    int64_t naccept, nreject, nrhs;
    int64_t nsaved;
    int retcode;   /* 0 Success, 1 MaxIters (the step count reached maxiters) */
    double ms;     /* device time of the step loop */
} eqd_bp8_stats;

/* run_bp8(m; tspan=(t0, tf), reltol, saveat) (src/BP8.jl:442-470): Tsit5 with the
 * OrdinaryDiffEq defaults (PI control beta1 = 7/50, beta2 = 2/25, qmin 1/5, qmax 10,
 * gamma 9/10; Hairer initial dt; RMS error norm; abstol 1e-14 / 1e-10 / 1e-3 (/ 1e-3) per
 * block; tstops = [t_off]; save_everystep = false). States at t0, t0 + saveat, ... (and tf
 * when it is on the grid) are written to ts[nsaved] and us[nsaved * state_len] (host,
 * row per save; NULL keeps them on the device, see eqd_bp8_saved_device).
 * maxiters <= 0 means 100000 (OrdinaryDiffEq's default). */
eqd_status eqd_bp8_integrate(eqd_bp8* m, double t0, double tf, double saveat, double reltol, int64_t maxiters,
                             int64_t nsave_max, double* ts, double* us, eqd_bp8_stats* stats);
/* run_bp8 with an implicit-local integrator instead of Tsit5, for the stiff Peaceman-well
 * variant (no counterpart in the reference, which scopes it in PROGRESS.md:872-894 and
 * TODO.md:126-153: the stiff mode is the well node's slip against its own self-stiffness, local
 * to one node). The 5-stage L-stable SDIRK of order 4 (Hairer & Wanner, gamma = 1/4) with its
 * embedded order-3 error estimate; each stage equation is solved by Newton iterations whose
 * Jacobian is the per-node 3 x 3 block of (s2, s3, ln theta), the dense K coupling and the pore
 * pressure converging as a fixed point. Steps end on every save time, on t_off and on tf (saves
 * are step ends). Same arguments, tolerances (reltol; abstol 1e-14 / 1e-10 / 1e-3 per block) and
 * saves as eqd_bp8_integrate; stats->nrhs counts right-hand-side evaluations. */
eqd_status eqd_bp8_integrate_imex(eqd_bp8* m, double t0, double tf, double saveat, double reltol,
                                  int64_t maxiters, int64_t nsave_max, double* ts, double* us,
                                  eqd_bp8_stats* stats);
/* How a model built with stiffness = TOEPLITZ forms K s (src/FaultResponse.jl:378-390 makes each of
 * K's four blocks a function of node separation, so K s is a 2D correlation with the
 * 4 (2 n2f - 1)^2-value separation kernel):
 *   EQD_KS_DIRECT  the correlation summed directly, O(nf^2) per product (inside the stage kernel
 *                  up to 48 x 48 fault nodes, the default there);
 *   EQD_KS_FFT     the same correlation as a zero-padded 2D FFT convolution (transform size the
 *                  power of two >= 2 n2f - 1), O(nf log nf); the default beyond 48 x 48 nodes
 *                  (fault patches of up to 256 x 256 nodes).
 * Both use K's own entries; results differ by rounding only. EQD_ERR_ARG for a dense-K model. */
enum { EQD_KS_DIRECT = 0, EQD_KS_FFT = 1 };
eqd_status eqd_bp8_set_ks_method(eqd_bp8* m, int method);
/* K s with the model's current method for a host slip vector s[2nf] = [s2; s3] into
 * out[2nf] = [tau2; tau3] (host); EQD_ERR_ARG when the method forms K s inside the stage kernel
 * (EQD_KS_DIRECT at 48 x 48 nodes or fewer). */
eqd_status eqd_bp8_ks(eqd_bp8* m, const double* s, double* out);
/* Measurement hook: average device time of one K s (the model's current method) over reps
 * launches, on the model's initial state. */
eqd_status eqd_bp8_ks_timing(eqd_bp8* m, int reps, double* ms_per_product);
/* Device pointer to the saved states of the last integrate: row j (save j) starts at
 * us_dev + j * ld (ld >= state_len, even). */
eqd_status eqd_bp8_saved_device(const eqd_bp8* m, const double** us_dev, int64_t* ld, int64_t* nsaved);
/* max_i |V_i| at every saved state of the last integrate, each evaluated with evaluate!
 * in save order as write_outputs does (src/BP8.jl:628). */
eqd_status eqd_bp8_saved_vmax(eqd_bp8* m, double* vmax, int64_t n);
/* The derived fields write_outputs computes at every saved state (src/BP8.jl:617-630): the
 * first n saves of the last integrate are evaluated with evaluate! in save order. For each save
 * j, out_global[2j] = max_i |V_i| and out_global[2j+1] = mu sum_i w_i |V_i| (moment rate, w the
 * Omega_f quadrature weights); for each listed Omega_f node k = idx[i] (0-based, x2 fastest),
 * out_fields[(j nidx + i) 6 + f] holds V2, V3, tau2, tau3, darcy_vel_2, darcy_vel_3 with
 * q_j = -(k / viscosity) D1_j p (src/PorePressure.jl:52-56). Either output may be NULL. */
eqd_status eqd_bp8_saved_outputs(eqd_bp8* m, int64_t n, const int* idx, int nidx, double* out_fields,
                                 double* out_global);

/* ------------------------------------------------------------------------------
 * Trajectory ensemble: M copies of run_bp8 (src/BP8.jl:442-470) on one model's K, A_p and
 * source, each member with its own Q0, a, sigma0 (and reltol), its own Tsit5 step
 * controller and saves. Members advance in lockstep stages; each stage's K S over all
 * members is one FP64 MFMA GEMM (K resident, S = the members' slip vectors), followed by
 * the per-(node, member) Newton in ln V, aging law and A_p p. A member's trajectory is the
 * one run_bp8 gives for build_model with that member's parameters (no counterpart in the
 * reference, which runs one trajectory per call). For a model built with stiffness = TOEPLITZ
 * the stage's K S is instead one FFT convolution batched over the members (EQD_KS_FFT).
 * ------------------------------------------------------------------------------ */
typedef struct eqd_bp8_ensemble eqd_bp8_ensemble;

/* Q0, a, sigma0: arrays of `members` values (NULL = the model's value for every member).
 * The model must outlive the ensemble. */
eqd_status eqd_bp8_ensemble_create(eqd_bp8* model, int members, const double* Q0, const double* a,
                                   const double* sigma0, eqd_bp8_ensemble** out);
eqd_status eqd_bp8_ensemble_destroy(eqd_bp8_ensemble* e);
/* Integrates every member over (t0, tf) as eqd_bp8_integrate does, with reltol for all or
 * member_reltol[members] (NULL = reltol). stats[members] (may be NULL); retcode 2 marks a
 * member whose Newton solve failed (the call then returns EQD_ERR_NEWTON). loop_ms: device
 * time of the step loop (may be NULL). */
eqd_status eqd_bp8_ensemble_integrate(eqd_bp8_ensemble* e, double t0, double tf, double saveat, double reltol,
                                      const double* member_reltol, int64_t maxiters, int64_t nsave_max,
                                      eqd_bp8_stats* stats, double* loop_ms);
/* Every member over (t0, tf) with the implicit-local integrator, as eqd_bp8_integrate_imex does
 * for one trajectory (the stiff Peaceman-well variant, src/BP8.jl:403-433 with the well of
 * src/PorePressure.jl:125-158): each member keeps its own step size, Newton iterations, error
 * control and saves; each Newton iteration forms K U and rhs! over all members at once. Same
 * arguments, stats and errors as eqd_bp8_ensemble_integrate (retcode 2: a member's stage solve
 * failed at the minimum step). No counterpart in the reference, which integrates with Tsit5. */
eqd_status eqd_bp8_ensemble_integrate_imex(eqd_bp8_ensemble* e, double t0, double tf, double saveat,
                                           double reltol, const double* member_reltol, int64_t maxiters,
                                           int64_t nsave_max, eqd_bp8_stats* stats, double* loop_ms);
/* A member's save times ts[nsaved] and states us[nsaved * state_len] (host; either may be NULL). */
eqd_status eqd_bp8_ensemble_saved(const eqd_bp8_ensemble* e, int member, double* ts, double* us,
                                  int64_t* nsaved);
/* max_i |V_i| at the first n saves of every member (vmax[member * n + j]), each save
 * evaluated in save order as eqd_bp8_saved_vmax does; one GEMM per save over all members.
 * ms: device time (may be NULL). */
eqd_status eqd_bp8_ensemble_saved_vmax(eqd_bp8_ensemble* e, double* vmax, int64_t n, double* ms);
/* Measurement hook: average device time of one K S over all members' current states (reps
 * launches) and the flop count 2 (2nf)^2 members of the GEMM. A member ensemble of a TOEPLITZ
 * model forms K S as one batched FFT convolution (no copy of K; flops reported as 0). */
eqd_status eqd_bp8_ensemble_gemm_timing(eqd_bp8_ensemble* e, int reps, double* ms_per_gemm, double* flops);

/* ------------------------------------------------------------------------------
 * Serving loop (no counterpart in the reference, whose K build is a batch job): stiffness
 * column and forward-solve requests on several grids, admitted into resident 64-column
 * tile pools as lanes free up (continuous batching), each column's result independent of
 * the mix it ran in.
 * ------------------------------------------------------------------------------ */
typedef struct eqd_server eqd_server;

/* opts: CG options of every grid (batched engine; tile_k is forced to 64). hbm_budget_bytes
 * bounds the tile pools of all grids together (0 = half of the free device memory). */
eqd_status eqd_server_create(const eqd_cg_opts* opts, size_t hbm_budget_bytes, eqd_server** out);
eqd_status eqd_server_destroy(eqd_server* sv);
/* Makes a grid resident: operator tables, a pool of 64-column tiles (as many as the
 * remaining budget holds, at most ceil(max_columns / 64) when max_columns > 0). On a grid
 * where one tile's launches already fill the device (dz = 25 m, L = (1600, 800)), the pool's
 * columns/s does not grow with its size while every column's latency does (each iteration
 * covers every active tile): give such a grid the lanes its largest request needs. */
eqd_status eqd_server_add_grid(eqd_server* sv, const eqd_grid_desc* g, int max_columns, int* grid_id, int* slots);
/* Queues a request for the unit-slip columns cols[0..ncols) (0-based, < 2nf) of K, i.e.
 * fault_stiffness restricted to those columns (src/FaultResponse.jl:229-287). */
eqd_status eqd_server_submit_columns(eqd_server* sv, int grid, const int* cols, int ncols, int64_t* req_id);
/* Queues a forward solve: shear_traction(fe, s2, s3) (src/FaultResponse.jl:194-211). */
eqd_status eqd_server_submit_forward(eqd_server* sv, int grid, const double* s2, const double* s3, int64_t* req_id);
/* One step: retire finished columns, admit queued requests into free lanes, run one graph
 * replay of CG iterations on every grid with work. Returns the lanes in flight and the
 * requests still queued. */
eqd_status eqd_server_step(eqd_server* sv, int64_t* lanes_busy, int64_t* queued);
/* Pops up to max completed request ids. */
eqd_status eqd_server_completed(eqd_server* sv, int64_t* ids, int max, int* n);
/* Result of a completed request (then forgotten): 2nf x ncols column-major [dtau2; dtau3]
 * (ld >= 2nf), CG iterations per column, count of columns that did not converge. */
eqd_status eqd_server_result(eqd_server* sv, int64_t id, double* out, int64_t ld, int* iters, int* unconverged);
/* Operator info, lane count and device bytes of a grid's pool. */
eqd_status eqd_server_grid_info(const eqd_server* sv, int grid, eqd_op_info* info, int* slots, double* hbm_bytes);

/* Version and last error. */
const char* eqd_version(void);
size_t eqd_last_error(char* buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif
