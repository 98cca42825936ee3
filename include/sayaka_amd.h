/* sayaka_amd.h -- C ABI of the MI355X (gfx950) MAC Poisson solve.
 *
 * This is the thin layer between the host-side tree/solver objects and the HIP kernels:
 * plain pointers and sizes, int status codes, no C++ types.  The reference has no device
 * and no FFI; each entry point names the reference MacSolver member (or step) whose
 * work it performs on the device.  The host C++ MacSolver (and the Python mirror) call
 * these after every regrid: topology -> BC -> coefficients -> right-hand side -> solve.
 *
 * Data layout on the host side of every call:
 *   cell field  [block][k][j][i], nb^d values per block (valid cells only)
 *   face field  dir d: [block][k][j][i] over the cell box staggered in d
 * Field ids follow the reference solver buffer components (SayakaMacSolver.h:34-53).
 *
 * Status: 0 ok; 1 solver breakdown; 2 not converged; negative = argument / HIP error
 * (text via sayaka_dev_last_error).  Nothing here exits or throws.
 * Threading: one context per host thread; every call is stream-ordered on the context's
 * stream and returns after the device work it needs has completed.
 */
#ifndef SAYAKA_AMD_H
#define SAYAKA_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sayaka_dev_ctx sayaka_dev_ctx;

enum sayaka_field {
    SAYAKA_RHS = 0, SAYAKA_SOL, SAYAKA_RESID, SAYAKA_CORR, SAYAKA_TMP, SAYAKA_X0, SAYAKA_X,
    SAYAKA_R, SAYAKA_RHAT, SAYAKA_P, SAYAKA_PHAT, SAYAKA_S, SAYAKA_SHAT, SAYAKA_V, SAYAKA_T,
    SAYAKA_NFIELD
};

enum sayaka_precision {
    SAYAKA_FP64 = 0,      /* every kernel in FP64, V-cycle and summation order exactly as the
                             reference: bit-identical results */
    SAYAKA_MIXED = 1,     /* FP32 V(2,2)-cycle inside the FP64 BiCGSTAB, FP64 coarsest solve */
    SAYAKA_FP64_FAST = 2  /* SAYAKA_FP64 with order-free (tree) sums in the inner products:
                             same algorithm, results differ from the reference by rounding */
};

typedef struct {
    double tol_rel;    /* MacSolver::setTolerance (SayakaMacSolver.h:350) */
    double tol_abs;
    int max_iter;      /* MacSolver::setMaxIteration (SayakaMacSolver.h:349) */
    int precond;       /* MacSolver::setCGUseMGPrecond (SayakaMacSolver.h:358) */
    int precision;     /* sayaka_precision */
    int pre_smooth;    /* mixed mode only (the reference V-cycle has none) */
    int post_smooth;   /* MacSolver::setPostSmoothNum (SayakaMacSolver.h:353), default 2 */
    int hist_cap;      /* capacity of stats->hist */
} sayaka_solve_params;

typedef struct {
    int status;        /* 0 converged, 1 breakdown, 2 max_iter reached */
    int iters;         /* outer BiCGSTAB iterations (the reference's "iter=%d" count) */
    int half_step;     /* 1 when convergence was detected on |s| */
    int bottom_calls;
    long long bottom_iters;
    double rnorm0, rnorm;  /* Linf of initial / final recursive residual */
    int nhist;
    double* hist;      /* caller-owned, hist_cap entries: -|s| (half) and +|r| (full) */
} sayaka_solve_stats;

const char* sayaka_dev_last_error(void);

/* MacSolver(tree) + MacSolver::define (SayakaMacSolver.h:56-87, SayakaMacOperator.cpp:13-60):
 * device buffers for up to max_blocks blocks of nb^ndim cells (nb = 8 or 16 in 2D / 3D, or
 * 4 in 3D). */
int sayaka_dev_create(int device, int ndim, int nb, int max_blocks, sayaka_dev_ctx** out);
int sayaka_dev_destroy(sayaka_dev_ctx* ctx);

/* MacSolver::prepareSolver (SayakaMacSolver.cpp:12-26): registerMGLevels from the tree
 * (SayakaMacOperator.cpp:62-78, SayakaTower.cpp:28-187) and unit coefficients.  The tree is
 * passed as its per-node arrays (AmrTreeNode, SayakaTree.h:52-82). */
int sayaka_dev_set_topology(sayaka_dev_ctx* ctx, int nblk, int minlevel, int maxlevel, const int* level,
                            const int* parent, const int* which_child, const int* child8, const int* nbr6,
                            const int* node_type, const double* length3);

/* MacSolver::setBC (SayakaMacOperator.cpp:1462-1469): per block face math BC type
 * (0 Neumann, 1 Dirichlet, 2 SimpleFill) on wall faces, -1 elsewhere; [nblk][6]. */
int sayaka_dev_set_bc(sayaka_dev_ctx* ctx, const int* bc_type6);

/* MacSolver::setScalars + setIBFracByLS/fractions + setCoefB (SayakaMacSolver.h:346,
 * SayakaMacIB.cpp:34-97, SayakaMacCoef.cpp:319-367): leaf values of the volume fraction,
 * area fractions and B per direction (NULL = unit); the device synchronises coarse/fine
 * faces, averages down to parent blocks and precomputes the face weights. */
int sayaka_dev_set_coefs(sayaka_dev_ctx* ctx, double alpha, double beta, const double* vf,
                         const double* const* af, const double* const* b);
/* read back the solver's coefficient state (tests): any pointer may be NULL;
 * w is [nblk][ndim][faces], wcf [nblk][2*ndim][nb^(ndim-1)] */
int sayaka_dev_get_coefs(sayaka_dev_ctx* ctx, double* vf, double* const* af, double* const* b, double* w,
                         double* wcf);

/* MacSolver::setRhs / setInitSol / getSol (SayakaMacSolver.h:117-132) on any field id */
int sayaka_dev_upload(sayaka_dev_ctx* ctx, int field, const double* host_valid);
int sayaka_dev_download(sayaka_dev_ctx* ctx, int field, double* host_valid);

/* MacSolver::fixBC (SayakaMacOperator.cpp:1470-1594): phi carries BC values in ngrow ghost
 * layers ([block] over the grown box); adds the wall corrections to field rhs_field. */
int sayaka_dev_fix_bc(sayaka_dev_ctx* ctx, const double* phi_grown, int ngrow, int rhs_field);

/* Operators on MG tower m (tests and fine-grained callers) */
int sayaka_dev_apply(sayaka_dev_ctx* ctx, int m, int out, int in);               /* mg_apply_level */
int sayaka_dev_residual(sayaka_dev_ctx* ctx, int m, int r, int x, int rhs);      /* mg_residual_level */
int sayaka_dev_smooth(sayaka_dev_ctx* ctx, int m, int sol, int rhs);             /* mg_smooth_level */
int sayaka_dev_restrict(sayaka_dev_ctx* ctx, int crse_level, int field);         /* mg_restict_wgt */
int sayaka_dev_prolong(sayaka_dev_ctx* ctx, int fine_level, int field);          /* mg_prolong_wgt */
int sayaka_dev_dot(sayaka_dev_ctx* ctx, int m, int a, int b, double* dot, double* maxabs_a);
                                                     /* mg_dotprod_level, mg_norm_level */
int sayaka_dev_precond(sayaka_dev_ctx* ctx, int m, int sol, int rhs);            /* mg_precond_level */
/* sayaka_dev_dot without any allocation.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops).  a and b
 * are device arrays [block][cell] in the context's layout (e.g. from sayaka_dev_alloc at set-up);
 * out2 is a device array of 2 doubles that receives dot(a, b) over tower m (summed in the same
 * order as sayaka_dev_dot, so the result is bit-identical to it) and max|a|.  Stream-ordered on
 * the context's stream: it returns without waiting (sayaka_dev_sync, or a copy on that stream,
 * orders a read of out2).  The per-block partials live in a fixed file-scope device array of
 * 1.5 MiB per GPU, so towers of at most 65 536 blocks (larger: -1); every call on one GPU shares
 * that array, so calls that overlap on one GPU (other streams, other contexts) race. */
int sayaka_dev_dot_noalloc(sayaka_dev_ctx* ctx, int m, const double* a, const double* b, double* out2);
int sayaka_dev_add(sayaka_dev_ctx* ctx, int m, int dst, int src);                /* mg_correct_level:
                                                        dst += src on tower m (SayakaMacUtil.cpp:109-133) */
int sayaka_dev_bottom(sayaka_dev_ctx* ctx, int sol, int rhs, double tol_rel, double tol_abs, int max_iter,
                      int* iters, int* status);      /* mg_solve_bottom */

/* MacSolver::mg_flux_level (SayakaMacOperator.cpp:161-239): the face flux -d(phi)/dx of
 * direction dir on tower m from the field phi, into flux_out [block][face] (host memory,
 * nblk * (nb+1)*nb^(ndim-1) doubles; blocks outside tower m are 0). */
int sayaka_dev_flux(sayaka_dev_ctx* ctx, int m, int dir, int phi, double* flux_out);

/* The one-layer face ghosts that MacSolver::mg_flux_level leaves in its phi argument on tower m
 * (its mg_fillbndry_level, SayakaMacOperator.cpp:181-182, SayakaMacUtil.cpp:210-251: injection, no
 * corners, no physical boundary) for the field: ghost_out (host) [block][face][slab], slab = the
 * face's nb^(ndim-1) cells with the lower tangential axis fastest; every block of tree levels
 * 1..m except a level whose next tower sub-level is empty; wall faces and unvisited blocks 0. */
int sayaka_dev_solver_ghosts(sayaka_dev_ctx* ctx, int m, int field, double* ghost_out);

/* Regrid data movement of the device fields, after sayaka_dev_set_topology with the regridded
 * tree: block b takes the values block origin[b] had before (origin from sy_tree_regrid_origin;
 * nold = block count before the regrid), and blocks with origin -1 are filled from their
 * parents by injection, coarsest level first -- the reference's packing of registered state
 * data and AmrTree::initNewBlockDataAfterRegrid with the default injection
 * (SayakaRegrid.cpp:530-591, 987-1105). */
int sayaka_dev_remap(sayaka_dev_ctx* ctx, const int* origin, int nold, const int* fields, int nfields);

/* Device memory of the context's GPU for caller-held data such as a state with ghost layers
 * (no reference counterpart: the reference's TreeData lives in host memory).  memcpy copies
 * host -> device when to_device is non-zero, device -> host otherwise, and returns when done. */
int sayaka_dev_alloc(sayaka_dev_ctx* ctx, size_t bytes, void** out);
int sayaka_dev_free(sayaka_dev_ctx* ctx, void* ptr);
int sayaka_dev_memcpy(sayaka_dev_ctx* ctx, void* dst, const void* src, size_t bytes, int to_device);

/* The regrid's packing of a caller's state held on the device (SayakaRegrid.cpp:530-591, which
 * moves every registered state's blocks with the tree's): block slot b of dst takes slot
 * origin[b] of src for b < n, zeros where origin[b] == -1 and in slots n..nslot-1; both states
 * [ncomp][comp_stride], a block slot block_len values, comp_stride >= nslot * block_len; src and
 * dst distinct.  With sayaka_dev_state_remap's new-block fill it keeps a state resident on the
 * device across regrids. */
int sayaka_dev_block_move(sayaka_dev_ctx* ctx, const double* src, double* dst, int ncomp, long long comp_stride,
                          long long block_len, int n, int nslot, const int* origin);

/* Host <-> device traffic of a context since its creation (no reference counterpart; it lets a
 * caller check that a state kept on the device is not copied back and forth): out4[0] bytes of
 * state to the device and out4[1] to the host (sayaka_dev_memcpy, sayaka_dev_upload /
 * download), out4[2] topology uploads (sayaka_dev_set_topology), out4[3] surroundings uploads
 * (sayaka_dev_set_surroundings). */
int sayaka_dev_transfer_stats(sayaka_dev_ctx* ctx, long long* out4);

/* Placement of the coarsest-level solve's persistent launch (no reference counterpart; the
 * reference solves on one CPU thread).  one_xcd = 1 (the default): when its workgroups fit on
 * one XCD they run there, and the grid barriers use that XCD's L2 as their coherence point
 * (no L2 write-back or invalidation); 0: spread over the device with device-wide barriers.
 * The results are identical either way. */
int sayaka_dev_set_bottom_placement(sayaka_dev_ctx* ctx, int one_xcd);

/* When sayaka_dev_solve queues a half-iteration before it has read whether the solve stops
 * (no reference counterpart; the reference tests on its one thread before going on,
 * SayakaMacSolver.cpp:683-920): with a preconditioner, whenever the last tested norm is at least
 * factor x the stopping threshold max(rnorm0 * tol_rel, tol_abs), the second half of an
 * iteration goes in before the host reads |s|, and (FP64 modes) the next iteration's first half
 * before it reads |r|.  The device applies the stopping tests itself; a queued V-cycle's
 * coarsest-level solve does nothing when they hold, a queued half step leaves x as it is, and
 * the other kernels of that half run and are discarded.  factor 0 queues every half, a negative
 * factor none.  Default 100.  Iterations, history, solution and coarsest-solve counts are
 * identical for every factor; only the time differs. */
int sayaka_dev_set_half_step_defer(sayaka_dev_ctx* ctx, double factor);

/* Coarsest-solve launches since the context's creation: out3[0] all, out3[1] placed on one XCD,
 * out3[2] of those, launches that found their workgroups on more than one XCD and kept the
 * device-wide barriers. */
int sayaka_dev_bottom_placement(sayaka_dev_ctx* ctx, long long* out3);

/* The tree's 3^d surrounding blocks for the general state fill, after every
 * sayaka_dev_set_topology: surr27[b][27] = SurroundingBlocks::blocks and bcref27[b][27] =
 * SurroundingBlocks::bcrefs of block b, both in the reference's [ii+1][jj+1][kk+1] order
 * (AmrTree::surrBlocks, SayakaTree.h:336-526, built by SayakaRegrid.cpp:655-945; block id >= 0
 * same level, -1 coarser, <= -20 physical code). */
int sayaka_dev_set_surroundings(sayaka_dev_ctx* ctx, int nblk, const int* surr27, const int* bcref27);

/* FillPatch::fillBoundary(scomp, nc, nlayer) (SayakaFillPatch.cpp:17-368) on a cell state held
 * on the device (state: device pointer, [ncomp][blocks_alloc][grown box], the block grown by
 * ngrow layers, i fastest, 2D blocks not grown in z): levels coarsest first, every block; same-
 * level surroundings (faces, edges, corners) copied, coarse ones prolonged from the parent
 * (prolong[c] per filled component: 0 injection, 1 centred linear, 2 MC-limited centred linear;
 * InterpPatch::ProlongMethod, SayakaInterpPatch.cpp:157-391), physical ones filled by the BC of
 * BCPatch::fillBasicCellBlockBC (SayakaBoundaryPatch.cpp:139-229) over all ngrow layers.  The
 * BC register of component scomp+c maps physical code physbc[h] at surrounding q to
 * (mathbc, bcval)[(c*27 + q)*nphys + h] (BCRegister, SayakaBoundaryPatch.h:62-140; math BC 0
 * Neumann, 1 Dirichlet, 2 SimpleFill).  Returns -2 when a physical code of the tree has no
 * mapping (the reference exits).  Needs sayaka_dev_set_surroundings for the current tree. */
int sayaka_dev_state_fill(sayaka_dev_ctx* ctx, double* state, int ncomp, int blocks_alloc, int ngrow, int scomp,
                          int nc, int nlayer, const int* prolong, int nphys, const int* physbc, const int* mathbc,
                          const double* bcval);

/* FillPatch::fillBoundary(scomp, nc, nlayer) on face data of direction fdir (0 x, 1 y, 2 z),
 * held on the device as [ncomp][blocks_alloc][grown face box]: the face box of a block has
 * nb+1 faces along fdir and nb across, grown by ngrow on every side (2D not in z), i fastest.
 * Same-level surroundings copied over the staggered ranges (where two ranges share a face the
 * surrounding later in the reference's loop order wins; FillPatch::cacheRanges,
 * SayakaFillPatch.cpp:620-700), coarse ones prolonged one after the other over their non-
 * overlapping ranges (prolongFaceBlockData_Injection / _CenteredLinear with or without the MC
 * limiter, SayakaInterpPatch.cpp:58-85, 447-591).  Valid faces are never written; the result
 * is the reference's for face data whose copies of a face shared by two blocks agree.  Returns
 * -2 on a tree with physical boundaries: the reference fills no face data next to a wall
 * (SayakaFillPatch.cpp:318-324).  Needs sayaka_dev_set_surroundings for the current tree. */
int sayaka_dev_face_fill(sayaka_dev_ctx* ctx, double* state, int ncomp, int blocks_alloc, int fdir, int ngrow,
                         int scomp, int nc, int nlayer, const int* prolong);

/* AmrTree::regrid's packing of a registered state and initNewBlockDataAfterRegrid
 * (SayakaRegrid.cpp:530-591, 987-1105) on the device, after sayaka_dev_set_topology and
 * sayaka_dev_set_surroundings for the NEW tree.  old_state: the state on the old tree
 * ([ncomp][old_blocks_alloc][grown box], nold blocks); state: the new state, a distinct buffer
 * ([ncomp][blocks_alloc][grown box]); fdir -1 for cell data, else the face direction (grown
 * face box, as for sayaka_dev_face_fill); origin[nblk]: each new block's id before the regrid,
 * -1 for a block the regrid created (sy_tree_regrid_origin).  Surviving blocks move whole; new
 * blocks start at 0; then, level by level from the coarsest holding a new block, the parents of
 * the level's new blocks are filled in-level over all ngrow layers with the physical BC
 * (fillBlockBoundary(.., onlySameLevel = 1); BC register as for sayaka_dev_state_fill), each
 * new block is prolonged over its valid box from its parent (prolongBlock; prolong[c] per
 * component) and, for face data, takes the shared faces of its existing same-level face
 * neighbours (correctFaceBlockData, SayakaFillPatch.cpp:485-513).  Face data: periodic trees
 * only (-2 otherwise, where the reference exits). */
int sayaka_dev_state_remap(sayaka_dev_ctx* ctx, const double* old_state, int old_blocks_alloc, int nold,
                           double* state, int ncomp, int blocks_alloc, int ngrow, int fdir, const int* origin,
                           const int* prolong, int nphys, const int* physbc, const int* mathbc, const double* bcval);

/* The valid cells of a device cell state ([ncomp][blocks_alloc][grown box], as for
 * sayaka_dev_state_fill) in the order of WriteDataHdf5's data:datatype=0 datasets
 * (SayakaWriterHdf5.cpp:420-512): for every tree level 1..finest, every block of the level in
 * ascending id, every component, the block's valid cells (i fastest).  out: device buffer of
 * ncomp * nblk * nb^d doubles; one copy to the host then feeds sy_write_chombo. */
int sayaka_dev_pack_levels(sayaka_dev_ctx* ctx, const double* state, int ncomp, int blocks_alloc, int ngrow,
                           double* out);

/* MacSolver::cg_solve (SayakaMacSolver.cpp:455-680): BiCGSTAB on the finest tower from
 * the SOL field as initial guess; SOL receives the solution on convergence. */
int sayaka_dev_solve(sayaka_dev_ctx* ctx, const sayaka_solve_params* p, sayaka_solve_stats* st);

/* Coarsest-level solve of the mixed-precision V-cycle (no reference counterpart: the FP64
 * mode always uses the reference's 1e-6 relative + 1e-6 absolute, 2000 iterations). */
int sayaka_dev_set_mixed_bottom(sayaka_dev_ctx* ctx, double tol_rel, double tol_abs, int max_iter);

/* Two-level preconditioner (Jacobi + one coarse unknown per level-1 block) for the mixed
 * solve's coarsest BiCGSTAB: 1 = on (the default; used when the level-1 problem has at
 * most 64 blocks), 0 = plain BiCGSTAB as in the FP64 mode.  No reference counterpart. */
int sayaka_dev_set_mixed_coarse(sayaka_dev_ctx* ctx, int two_level);

/* Smoother of the mixed solve's V-cycle: 2 = hybrid sweeps, one launch per sweep over the
 * whole tower, each block red then black in LDS with its halo as it was before the launch,
 * with post-sweeps beyond sayaka_solve_params.post_smooth (one more on 16^3 blocks, three more
 * on smaller ones) for the stale block boundaries (the default); 1 = exact red-black, each
 * colour in one launch over every sub-level of the tower, double-buffered; 0 = exact
 * red-black, one launch per sub-level and colour in place, in the reference's sub-level
 * order.  No reference counterpart. */
int sayaka_dev_set_mixed_smoother(sayaka_dev_ctx* ctx, int merged);

/* Copy one field into another on the device (TreeData::Copy of the valid region,
 * SayakaTreeData.cpp:267-290). */
int sayaka_dev_copy(sayaka_dev_ctx* ctx, int dst, int src);

/* Set every value of a field to zero on the device (the reference's setZeroSol,
 * SayakaMacSolver.h:125-128, without a host copy). */
int sayaka_dev_fill(sayaka_dev_ctx* ctx, int field, double value);

/* Launch timing per kernel class, with HIP events on the context stream (no reference
 * counterpart; used by the benchmark).  enable != 0 clears and starts recording; reading
 * returns the summed event time, the launch count, the cells covered and the algorithmic
 * HBM bytes (each value the launches must read or write once; 0 for the bottom solve). */
enum sayaka_kernel_class {
    SAYAKA_KCLASS_APPLY = 0,     /* FP64 operator / residual stencils (fused dot products) */
    SAYAKA_KCLASS_SMOOTH,        /* FP64 red-black Gauss-Seidel colour sweeps */
    SAYAKA_KCLASS_SMOOTH32,      /* FP32 red-black sweeps of the mixed V-cycle */
    SAYAKA_KCLASS_BOTTOM,        /* coarsest-level BiCGSTAB (one persistent launch) */
    SAYAKA_KCLASS_VEC,           /* fused vector updates and reductions */
    SAYAKA_KCLASS_TRANSFER,      /* restriction / prolongation */
    SAYAKA_KCLASS_N
};
int sayaka_dev_profile(sayaka_dev_ctx* ctx, int kclass, int enable);
int sayaka_dev_profile_read(sayaka_dev_ctx* ctx, int kclass, double* total_ms, long long* launches,
                            long long* cells, double* alg_bytes);

/* ---- learned initial guess (no reference counterpart) ---------------------------------
 * A per-block 3D CNN (12^3 window: the block and two same-level ghost cells per face; 5 input
 * channels: scaled rhs, three cell-averaged B*a, volume fraction; four 3x3x3 convolutions
 * 16-16-16-1 with ReLU; BF16 weights and activations, FP32 sums on the matrix cores) that maps
 * each leaf block to an initial guess of its cells.  8^3 blocks in 3D only.
 * Parameters: 4 layers x (27*16*16 weights [tap][ci][co] + 16 biases), FP32. */
int sayaka_cnn_setup(sayaka_dev_ctx* ctx, int max_windows);
int sayaka_cnn_set_params(sayaka_dev_ctx* ctx, const float* params);
int sayaka_cnn_get_params(sayaka_dev_ctx* ctx, float* params);
/* training store: windows of the current tree's leaves (ids on the host) with targets
 * (SOL - shift) * t_scale; collect returns the store size */
int sayaka_cnn_store_reset(sayaka_dev_ctx* ctx, long long capacity);
long long sayaka_cnn_collect(sayaka_dev_ctx* ctx, const int* leaf_ids, int n, double s_rhs, double shift,
                             double t_scale);
int sayaka_cnn_store_read(sayaka_dev_ctx* ctx, long long first, int n, float* windows, float* targets);
/* one Adam step (lr) on the MSE of the store windows idx[n]; loss of the batch before the step */
int sayaka_cnn_train_step(sayaka_dev_ctx* ctx, const long long* idx, int n, float lr, float* loss);
/* loss, parameter gradients and (optional) FP32 outputs [n][12^3] without an update */
int sayaka_cnn_loss_grad(sayaka_dev_ctx* ctx, const long long* idx, int n, float* loss, float* grad, float* out);
/* initial guess of every leaf block into `field` (output * out_scale; solid cells and
 * non-leaf blocks zero) */
int sayaka_cnn_guess(sayaka_dev_ctx* ctx, int field, double s_rhs, double out_scale);
/* Inputs of the learned-guess workload for the current tree, generated on the device: solid
 * sphere (centre, radius) with cut-cell fractions (nsub samples per edge), B = 1/rho
 * (rho = 1 + (ratio-1) H_eps(-phi)), rhs = sum of nblob Gaussians (centres [nblob][3], widths,
 * amplitudes) on fluid cells projected to zero mean (the removed mean in *mean_out).
 * lo3 [nblk][3] and dx [nblk] describe the blocks. */
int sayaka_gen_sphere_blobs(sayaka_dev_ctx* ctx, const double* lo3, const double* dx, const double* centre,
                            double radius, double ratio, double eps, int nsub, int nblob, const double* blob_c,
                            const double* blob_w, const double* blob_a, double* mean_out);
/* The sphere's coefficients of sayaka_gen_sphere_blobs alone (vf, af, B; coarse/fine faces
 * synchronised), leaving the rhs as it is. */
int sayaka_gen_sphere(sayaka_dev_ctx* ctx, const double* lo3, const double* dx, const double* centre, double radius,
                      double ratio, double eps, int nsub);

/* Projection of a face velocity, on the device (the projection step of the reference's flow
 * demo, Test7.h:580-657, with dt = 1 and p's sign flipped: rhs = div u, solve, u -= B flux with
 * mg_flux_level's flux,
 * fill_vel_all's coarse/fine synchronisation).  No single reference function: the demo drives
 * TreeData face fields and MacSolver from the host.
 *
 * sayaka_gen_velocity: a seeded face velocity u ~ U(-1,1) for the current tree, keyed on (seed,
 * level, global face index, direction) so that a face shared by two blocks of one level gets one
 * value (lo3 [nblk][3] block corners, dlo3 the domain's lower corner, root3 the root block edge);
 * zero_walls = 1 sets it to 0 on the domain walls.  Invalidated by sayaka_dev_set_topology.
 * sayaka_gen_divergence_rhs: synchronises the velocity's coarse/fine faces (coarse side = average
 * of the fine faces) and writes rhs = div(a u) / (dv vf) on fluid cells (0 on solid cells) minus
 * its volume-weighted mean over the fluid leaf cells (returned in *mean_out).
 * sayaka_dev_project: u -= B flux(phi) on every face of the finest tower (flux as
 * sayaka_dev_flux), faces re-synchronised; *div_before / *div_after = max |div u - mean| over
 * the tower before and after.
 * sayaka_dev_get_velocity: the face velocity to the host, [nblk][ndim][(nb+1) nb^(ndim-1)]. */
int sayaka_gen_velocity(sayaka_dev_ctx* ctx, const double* lo3, const double* dlo3, const double* root3,
                        unsigned long long seed, int zero_walls);
int sayaka_gen_divergence_rhs(sayaka_dev_ctx* ctx, double* mean_out);
int sayaka_dev_project(sayaka_dev_ctx* ctx, int phi, double* div_before, double* div_after);
int sayaka_dev_get_velocity(sayaka_dev_ctx* ctx, double* out);

/* Block until the context's stream is idle. */
int sayaka_dev_sync(sayaka_dev_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
