/* CPU FP64 oracle for the split-node fault-stiffness path and the BP8 right-hand side.
 *
 * TEST INFRASTRUCTURE ONLY. This library restates the reference's CPU algorithm
 * (assembled CSR operators, Krylov-style CG, the K build, the BP8 RHS) so that the
 * HIP path can be checked against it. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker or the timed CPU
 * baseline. The product library (libeqd.so) never links or calls it.
 *
 * Reference functions restated (paths relative to the reference repository):
 *   eqo_create               src/ElasticitySplitNode.jl:106-214 (split_node_system),
 *                            src/FaultResponse.jl:96-146 (FaultElasticity),
 *                            src/Elasticity.jl:44-89,195-221 (elastic_blocks, traction_blocks)
 *   eqo_cg                   src/ElasticitySplitNode.jl:378-396 -> Krylov.cg! 0.10.9 (x0 = 0,
 *                            stop when ||r|| <= atol + rtol*||b||, itmax = 0 -> 2n)
 *   eqo_shear_traction       src/FaultResponse.jl:194-211
 *   eqo_fault_stiffness      src/FaultResponse.jl:229-287 (strided threaded partition)
 *   eqo_fault_stiffness_toeplitz  src/FaultResponse.jl:339-393
 */
#ifndef EQD_ORACLE_H
#define EQD_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int n1;           /* nodes per block across the fault (x1) */
    int n23;          /* nodes per block along strike and depth (x2 = x3) */
    double L_normal;  /* block extent in x1 */
    double L_fault;   /* half-width in x2 and x3 */
    double lambda, mu;
    double l_f;       /* half-width of the frictional patch */
    int precond;      /* 0 = none, 1 = jacobi */
    double rtol, atol;
    int itmax;        /* 0 = 2 * size (Krylov's default) */
} eqo_params;

typedef struct {
    int64_t solves, iterations, max_iterations, unconverged;
} eqo_stats;

/* info indices */
enum { EQO_NF = 0, EQO_N2F, EQO_N3F, EQO_NTOT, EQO_NNZ_A, EQO_NNZ_HPDSAT, EQO_NNZ_P, EQO_N_PER_BLOCK,
       EQO_I2LO, EQO_I3LO, EQO_INFO_LEN };
/* operators for eqo_apply / eqo_export */
enum { EQO_OP_A = 0, EQO_OP_HPDSAT, EQO_OP_P, EQO_OP_T2, EQO_OP_T3, EQO_OP_DSAT, EQO_OP_H };

void* eqo_create(const eqo_params* p, char* err, size_t errlen);
void eqo_destroy(void* h);
int eqo_info(void* h, int64_t* out);
int eqo_apply(void* h, int op, const double* x, double* y);
int64_t eqo_export(void* h, int op, int64_t* rowptr, int64_t* col, double* val); /* returns nnz; NULL bufs = query */
int eqo_build_chi(void* h, const double* s2, const double* s3, double* chi);
int eqo_cg(void* h, const double* b, double* x, int* niter, int* solved);
int eqo_shear_traction(void* h, const double* s2, const double* s3, double* dtau2, double* dtau3);
int eqo_fault_stiffness(void* h, int nthreads, double* K);
int eqo_fault_stiffness_cols(void* h, const int* cols, int ncols, int nthreads, double* Kcols);
int eqo_fault_stiffness_toeplitz(void* h, double* K);
int eqo_report(void* h, eqo_stats* st);
void eqo_reset_stats(void* h);

/* 1D SBP order-4 operators for identity tests: H (n), D1 (n*n), D2 (n*n), dl (n), dr (n). */
int eqo_sbp1d(int n, double h, double* H, double* D1, double* D2, double* dl, double* dr);

#ifdef __cplusplus
}
#endif
#endif
