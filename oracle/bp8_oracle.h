/* CPU FP64 oracle of the BP8 time loop (right-hand side and Tsit5 driver).
 *
 * TEST INFRASTRUCTURE ONLY -- the same rules as eqd_oracle.h: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
 *
 * Reference functions restated:
 *   eqo_solve_slip_rate       src/RateStateFriction.jl:77-96 (Newton in ln V, clamp 5, tol 1e-12, 50 its)
 *   eqo_bp8_create            src/BP8.jl:184-258 (build_model minus the K build: A_p, source,
 *                             weights, active ring, tau0, Peaceman WI; src/PorePressure.jl:36-43,125-158)
 *   eqo_bp8_initial_state     src/BP8.jl:337-345, src/RateStateFriction.jl:127-130
 *   eqo_bp8_evaluate          src/BP8.jl:355-396 (mutates the warm-start cache like the reference)
 *   eqo_bp8_rhs               src/BP8.jl:403-433
 *   eqo_bp8_run               src/BP8.jl:442-470 with OrdinaryDiffEq 7.6 Tsit5 semantics: PI step
 *                             control (beta1 = 7/50, beta2 = 2/25, qmin 1/5, qmax 10, gamma 9/10,
 *                             qoldinit 1e-4), Hairer's initial dt, RMS error norm, tstops = [t_off],
 *                             saveat via the Tsit5 free interpolant, save_everystep = false.
 */
#ifndef BP8_ORACLE_H
#define BP8_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {  /* BP8Params, Table 1 (src/BP8.jl:50-87) */
    double mu, nu, rho, c_s;
    double a, b, D_RS, V_star, f_star;
    double sigma0, tau_init, V_init, V_zero;
    double Q0, L_gauss, L_fwid, beta, phi, k, viscosity, alpha, S_well, r_well;
    double l_f, dz, t_off, t_f, sigma_min;
} eqo_bp8_params;

void eqo_bp8_default_params(eqo_bp8_params* p);

/* Newton solve of T = eta V + sigma a asinh(V/(2V*) C) for V; status 0 ok, 1 no convergence. */
int eqo_solve_slip_rate(double T, double theta, double sigma, double eta, double a, double b, double Dc,
                        double V_star, double f_star, double V0, double tol, int maxiter, double maxstep, double* V);

/* injection: 0 Gaussian source, 1 Peaceman well. K is 2nf x 2nf column-major (copied). */
void* eqo_bp8_create(const eqo_bp8_params* p, const double* K, int n2f, int injection, char* err, size_t errlen);
void eqo_bp8_destroy(void* h);
int eqo_bp8_state_length(void* h);
int eqo_bp8_initial_state(void* h, double* u);
int eqo_bp8_evaluate(void* h, const double* u, double t, double* V2, double* V3, double* Vmag, double* tau2,
                     double* tau3);
int eqo_bp8_rhs(void* h, const double* u, double t, double* du);
/* A_p as a dense nf x nf row-major matrix (tests). */
int eqo_bp8_Ap(void* h, double* Ap);
int eqo_bp8_source(void* h, double* src, double* weights);
/* Integrates [t0, tf]; stores states at t0, t0 + saveat, ..., tf (nsave_max rows of state_length).
 * Also V_max (max |V| over Omega_f) at each saved time. Returns 0, or 1 on a Newton failure. */
int eqo_bp8_run(void* h, double t0, double tf, double saveat, double reltol, int nsave_max, double* ts,
                double* us, double* vmax, int* nsaved, int64_t* naccept, int64_t* nreject, int64_t* nrhs);
void eqo_bp8_reset_cache(void* h);

#ifdef __cplusplus
}
#endif
#endif
