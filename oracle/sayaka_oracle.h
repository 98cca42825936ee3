/* sayaka_oracle.h -- CPU restatement of the reference MAC Poisson solve (plain C).
 *
 * TEST INFRASTRUCTURE ONLY.  Used by tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py as the checker of the HIP path.  Never linked into the product.
 *
 * The context takes a tree as flat arrays (the export of sy_tree_export or of the
 * reference harness) and holds the solver state the reference MacSolver holds:
 * a 15-component buffer with one ghost layer (SayakaMacSolver.h:34-53), one face array
 * per direction for coarse/fine contributions, IB fractions, B coefficients and the
 * per-block BC records.  Every operation reproduces the reference's arithmetic in the
 * reference's order, so results are bit-identical to the compiled reference.
 *
 * Layouts (shared with the rest of the project):
 *   grown cell field : [block][k][j][i] over the block grown by one cell (z not grown in 2D)
 *   valid cell field : [block][k][j][i] over the nb^d valid cells
 *   face field dir d : [block][k][j][i] over the cell box staggered in d
 */
#ifndef SAYAKA_ORACLE_H
#define SAYAKA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oc_ctx oc_ctx;

enum { OC_RHS, OC_SOL, OC_RESID, OC_CORR, OC_TMP, OC_X0, OC_X, OC_R, OC_RHAT, OC_P, OC_PHAT,
       OC_S, OC_SHAT, OC_V, OC_T, OC_NCOMP };

oc_ctx* oc_create(int ndim, int nb, int nblk, int minlevel, int maxlevel, const int* level,
                  const int* parent, const int* which_child, const int* child8, const int* nbr6,
                  const int* node_type, const double* length3);
void oc_destroy(oc_ctx* c);

/* tables */
int oc_tower_size(const oc_ctx* c, int m);
void oc_tower_export(const oc_ctx* c, int m, int* blocks, int* begin, int* end, int* face_type6,
                     int* face_neigh6, int* fc_off18, int* cf_neigh24, int* cf_off72, int* cf_box144);

/* set-up (SayakaMacSolver.cpp:12-26, SayakaMacCoef.cpp, SayakaMacIB.cpp, SayakaMacOperator.cpp:1462-1594) */
void oc_prepare(oc_ctx* c);
void oc_set_coefb(oc_ctx* c, const double* const* b);
void oc_set_ib_frac(oc_ctx* c, const double* vf, const double* const* af);
void oc_set_ib_ls(oc_ctx* c, const double* ls_grown, int ngrow);
void oc_get_coefs(const oc_ctx* c, double* vf, double* const* b, double* const* af);
void oc_set_bc(oc_ctx* c, int n, const int* physbc, const int* mathbc, const double* val);
void oc_fix_bc(const oc_ctx* c, const double* phi_grown, int ngrow, double* rhs_valid);
void oc_set_params(oc_ctx* c, double tol_rel, double tol_abs, int max_iter, int precond);
void oc_set_rhs(oc_ctx* c, const double* rhs_valid);
void oc_set_init_sol(oc_ctx* c, const double* sol_valid);
void oc_set_zero_sol(oc_ctx* c);
void oc_get_sol(const oc_ctx* c, double* sol_valid);

/* buffer access (grown layout) */
void oc_buf_set(oc_ctx* c, int comp, const double* grown);
void oc_buf_get(const oc_ctx* c, int comp, double* grown);
void oc_save_get(const oc_ctx* c, int dir, double* faces);

/* operations on the buffer */
void oc_fill(oc_ctx* c, int m, int comp);
/* MacSolver::mg_flux_level (SayakaMacOperator.cpp:161-203, 929-1025): flux of direction d on
 * tower m from valid phi [blk][cell] into out [blk][face] */
void oc_flux(oc_ctx* c, int m, int d, const double* phi, double* out);
void oc_apply(oc_ctx* c, int m, int outcomp, int incomp);
void oc_residual(oc_ctx* c, int m, int rcomp, int xcomp, int rhscomp);
void oc_smooth(oc_ctx* c, int m, int solcomp, int rhscomp);
void oc_restrict(oc_ctx* c, int crse_level, int comp);
void oc_prolong(oc_ctx* c, int fine_level, int comp);
double oc_dot(const oc_ctx* c, int m, int a, int b);
double oc_norm(const oc_ctx* c, int m, int a);
int oc_precond(oc_ctx* c, int m, int solcomp, int rhscomp);

/* BiCGSTAB on one tower (SayakaMacSolver.cpp:683-920).  hist (may be NULL) receives -|s|
 * per half step and +|r| per full step, as the reference logs them. */
int oc_cg_level(oc_ctx* c, int m, int solcomp, int rhscomp, double tol_rel, double tol_abs,
                int max_iter, int precond, int* iters, int* converged, double* hist, int hist_cap,
                int* nhist);
/* MacSolver::cg_solve (SayakaMacSolver.cpp:455-680) */
int oc_cg_solve(oc_ctx* c, int* iters, int* converged, double* hist, int hist_cap, int* nhist);
/* counters of bottom solves since the last reset: calls and total iterations */
void oc_bottom_stats(oc_ctx* c, int reset, int* calls, long long* iters);

#ifdef __cplusplus
}
#endif
#endif
