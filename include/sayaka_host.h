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
 * bc[6]: neighbour code of each wall (<= -20).  Returns NULL on bad arguments. */
sy_tree* sy_tree_create(int ndim, int nb, int ngrow, int minlevel, int maxlevel, int max_blocks,
                        const int* ndiv, const double* lo, const double* hi, const int* bc);
void sy_tree_destroy(sy_tree* t);
/* AmrTree::numBlocks, AmrTree::currentFinestLevel (SayakaTree.h:201, 305-315) */
int sy_tree_numblocks(const sy_tree* t);
int sy_tree_finest(const sy_tree* t);
/* AmrTree::refineFlag / coarsenFlag [0, n) (SayakaTree.h:221-222); NULL leaves one unchanged */
int sy_tree_set_flags(sy_tree* t, const int* refine, const int* coarsen, int n);
/* AmrTree::regrid (SayakaRegrid.cpp:26-125).  Returns numBlocks, -1 on pool overflow. */
int sy_tree_regrid(sy_tree* t);
/* Flat copy of AmrTreeNode and SurroundingBlocks (SayakaTree.h:52-152, 336-526) */
void sy_tree_export(const sy_tree* t, int* level, int* parent, int* which_child, int* child8,
                    int* nbr6, int* node_type, int* node_type2, double* lo3, double* hi3,
                    double* center3, double* length3, int* surr27, int* surrtype27,
                    int* surrbcref27);
/* AmrTree::getTreeLevel (SayakaTree.h:317-318) */
int sy_tree_level_count(const sy_tree* t, int level);
void sy_tree_level_blocks(const sy_tree* t, int level, int* out);

#ifdef __cplusplus
}
#endif
#endif
