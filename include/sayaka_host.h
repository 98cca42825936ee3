/* sayaka_host.h -- C ABI of the host-side tree and solver driver.
 *
 * The reference has no FFI: its callers are C++ demo drivers that construct AmrTree and
 * MacSolver objects (Test4.h:186-213, Test5.h:231-263, Test6.h:254-296, Test7.h:411-418,
 * 612-653).  This header is the flat form of that C++ surface, for callers that are not
 * C++ (ctypes, cgo, JNI).  Every function names the reference member it stands for.
 * Handles are opaque; arrays are caller-owned; nothing here throws or exits.
 *
 * Layouts: cell fields [block][k][j][i] over the block's valid cells (nb^d values per
 * block, i fastest, SayakaBox.h:214-217); face fields of direction d [block][k][j][i] over
 * the cell box staggered in d ((nb+1)*nb^(d-1) values per block).
 */
#ifndef SAYAKA_HOST_H
#define SAYAKA_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sy_tree sy_tree;
typedef struct sy_mac sy_mac;

/* AmrTree::init + initUniformRootLevel + initRefineToMinLevel (SayakaTree.cpp:34-233).
 * bc[6]: neighbour code of each wall (<= -20), or SY_TREE_PERIODIC for a periodic face
 * (initUniformRootLevel's is_periodic, SayakaTree.cpp:93-201).  Returns NULL on bad arguments. */
#define SY_TREE_PERIODIC 1
sy_tree* sy_tree_create(int ndim, int nb, int ngrow, int minlevel, int maxlevel, int max_blocks,
                        const int* ndiv, const double* lo, const double* hi, const int* bc);
void sy_tree_destroy(sy_tree* t);
/* AmrTree::numBlocks, AmrTree::currentFinestLevel (SayakaTree.h:201, 305-315) */
int sy_tree_numblocks(const sy_tree* t);
int sy_tree_finest(const sy_tree* t);
/* AmrTree::refineFlag / coarsenFlag [0, n) (SayakaTree.h:221-222); NULL leaves one unchanged */
int sy_tree_set_flags(sy_tree* t, const int* refine, const int* coarsen, int n);
/* the current refineFlag / coarsenFlag of blocks [0, n) (either pointer may be NULL) */
int sy_tree_get_flags(const sy_tree* t, int* refine, int* coarsen, int n);
/* AmrTree::regrid (SayakaRegrid.cpp:26-125).  Returns numBlocks, -1 on pool overflow. */
int sy_tree_regrid(sy_tree* t);
/* after sy_tree_regrid: per block its id before the regrid, -1 for a block the regrid created
 * (the data movement of SayakaRegrid.cpp:530-591; no reference accessor, the reference moves its
 * registered state data itself) */
void sy_tree_regrid_origin(const sy_tree* t, int* origin);
/* Flat copy of AmrTreeNode and SurroundingBlocks (SayakaTree.h:52-152, 336-526) */
void sy_tree_export(const sy_tree* t, int* level, int* parent, int* which_child, int* child8,
                    int* nbr6, int* node_type, int* node_type2, double* lo3, double* hi3,
                    double* center3, double* length3, int* surr27, int* surrtype27,
                    int* surrbcref27);
/* Refine level by level around the shell |x - centre| = r0 (blocks with a cell centre within
 * factor * blockLength of it), as the demos' refinement loops do with their own criteria
 * (Test7.h:411-418).  Returns numBlocks, -1 on pool overflow. */
int sy_tree_build_shell(sy_tree* t, const double* centre, double r0, double factor);
/* refine / coarsen flags of the current leaves for the shell |x - centre| = r0 (the rule of
 * sy_tree_build_shell); returns the number of flagged blocks; follow with sy_tree_regrid */
int sy_tree_flag_shell(sy_tree* t, const double* centre, double r0, double factor);
/* States registered with the tree (TreeStateData, SayakaTree.h:528-560, as the demos register
 * theirs, Test5.h:590-610): cell data (dir < 0) or face data of direction dir, ncomp
 * components, ngrow ghost layers, a FillPatch with prolong[c] per component and, for cell data,
 * a BCPatch mapping each of the nphys physical codes to (mathbc, bcval)[c * nphys + q] on every
 * surrounding position.  Owned by the tree handle.  sy_tree_regrid moves their blocks
 * (SayakaRegrid.cpp:530-591); sy_tree_init_new_data is AmrTree::initNewBlockDataAfterRegrid
 * (SayakaRegrid.cpp:987-1105), run on the GPU.  Data as [comp][block][grown box] for the
 * tree's current blocks.  Returns: the state index / 0, or -1 (sy_tree_last_error). */
const char* sy_tree_last_error(void);
int sy_tree_state_register(sy_tree* t, int dir, int ncomp, int ngrow, const int* prolong, int nphys,
                           const int* physbc, const int* mathbc, const double* bcval);
int sy_tree_state_set_grown(sy_tree* t, int idx, const double* data);
int sy_tree_state_get_grown(const sy_tree* t, int idx, double* data);
int sy_tree_init_new_data(sy_tree* t);
/* FillPatch::fillBoundary(scomp, ncomp, nlayer) (SayakaFillPatch.cpp:17-30) of a registered
 * state, through its FillPatch */
int sy_tree_state_fill(sy_tree* t, int idx, int scomp, int ncomp, int nlayer);
/* TreeData::setResident (no reference counterpart): on != 0 keeps a copy of the state on GPU
 * `device`, where regrid moves, new-block fills and ghost fills then run with no host transfer;
 * set_grown / get_grown synchronise it; on == 0 brings it back and frees it. */
int sy_tree_state_set_resident(sy_tree* t, int idx, int on, int device);
/* host <-> device traffic of the tree's device context (sayaka_dev_transfer_stats) */
int sy_tree_transfer_stats(const sy_tree* t, long long* out4);

/* AmrTree::getTreeLevel (SayakaTree.h:317-318) */
int sy_tree_level_count(const sy_tree* t, int level);
void sy_tree_level_blocks(const sy_tree* t, int level, int* out);

/* ---- MacSolver (SayakaMacSolver.h:25-425) over a tree; solves on the GPU -------------
 * Integer returns: 0 ok (cg_solve: the reference's 0 / 1), < 0 error (sy_mac_last_error). */
const char* sy_mac_last_error(void);
/* MacSolver(const AmrTree&) (SayakaMacSolver.h:56): device context sized by the tree's pool */
sy_mac* sy_mac_create(sy_tree* tree, int device);
void sy_mac_destroy(sy_mac* m);
/* MacSolver::prepareSolver (SayakaMacSolver.cpp:12-26); call after every regrid */
int sy_mac_prepare(sy_mac* m);
/* MacSolver::setBC (SayakaMacOperator.cpp:1462-1469) with a BCRegister that maps each
 * phys_bc[q] to math_bc[q] (0 Neumann, 1 Dirichlet, 2 SimpleFill) on every face */
int sy_mac_set_bc(sy_mac* m, int n, const int* phys_bc, const int* math_bc);
/* MacSolver::setIBFracByLS (SayakaMacIB.cpp:12-154): level set over the grown box [blk][...] */
int sy_mac_set_ib_ls(sy_mac* m, const double* ls_grown, int ngrow);
/* cut-cell volume fractions [blk][cells] and area fractions af[dir][blk][faces] */
int sy_mac_set_ib_frac(sy_mac* m, const double* vf, const double* const* af);
/* MacSolver::setCoefB (SayakaMacCoef.cpp:319-367): b[dir][blk][faces] */
int sy_mac_set_coefb(sy_mac* m, const double* const* b);
/* MacSolver::fixBC (SayakaMacSolver.h:166, SayakaMacOperator.cpp:1470-1594): phi over the grown
 * box [blk][grown box] (ngrow >= 1 layers, i fastest; the wall ghosts carry the BC values), rhs
 * packed [blk][cells], updated in place with the Dirichlet / SimpleFill wall terms */
int sy_mac_fix_bc(sy_mac* m, const double* phi_grown, int ngrow, double* rhs);
/* MacSolver::setRhs / setInitSol / setZeroSol / getSol (SayakaMacSolver.h:117-132) */
int sy_mac_set_rhs(sy_mac* m, const double* rhs);
int sy_mac_set_init_sol(sy_mac* m, const double* sol);
int sy_mac_set_zero_sol(sy_mac* m);
int sy_mac_get_sol(sy_mac* m, double* sol);
/* mg_flux_level (SayakaMacSolver.h:275, SayakaMacOperator.cpp:161-239): face flux of
 * direction dir on tower mg_level from packed phi into packed flux [blk][face] */
int sy_mac_flux(sy_mac* m, int mg_level, int dir, const double* phi, double* flux);
/* setTolerance / setMaxIteration / setCGUseMGPrecond / setPre-PostSmoothNum
 * (SayakaMacSolver.h:349-358); precision 0 = FP64 (reference order),
 * 1 = mixed, 2 = FP64 with order-free sums */
int sy_mac_set_params(sy_mac* m, double tol_rel, double tol_abs, int max_iter, int precond, int precision,
                      int pre_smooth, int post_smooth);
/* MacSolver::cg_solve (SayakaMacSolver.cpp:455-680) */
int sy_mac_cg_solve(sy_mac* m, int* iters, int* status, double* rnorm0, double* rnorm);
/* MacSolver::mg_solve (SayakaMacSolver.cpp:43-162) with setMGFixedIter(fixed_iter) (0: iterate
 * to the tolerance, up to max_iter V-cycles) and MacSolver::relax_solve (377-452); from the
 * current solution; status 0 converged, 2 not (the reference returns 0 either way) */
int sy_mac_mg_solve(sy_mac* m, int fixed_iter, int* iters, int* status, double* rnorm0, double* rnorm);
int sy_mac_relax_solve(sy_mac* m, int* iters, int* status, double* rnorm0, double* rnorm);

/* ---- Chombo-format HDF5 output (WriteDataHdf5, SayakaWriterHdf5.cpp:351-530) -----------
 * Cell data already in the file's order, packed = for every tree level 1..finest, every block
 * of the level in ascending id, every component, the block's valid cells (i fastest) -- what
 * sayaka_dev_pack_levels produces.  WriterMeta (SayakaWriter.h:13-48): time, dt, step
 * (written as "iteration"), names[ncomp] (NULL, or NULL entries: "comp<i>").  Written by the
 * build's own HDF5 writer; no HDF5 library needed.  Returns 0, -1 on a bad argument or an I/O
 * error (sy_writer_last_error). */
const char* sy_writer_last_error(void);
int sy_write_chombo(const sy_tree* t, int ncomp, const double* packed, const char* filename, double time,
                    double dt, int step, const char* const* names);

#ifdef __cplusplus
}
#endif
#endif
