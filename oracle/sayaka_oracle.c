/* sayaka_oracle.c -- CPU restatement of the reference MAC Poisson solve.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle): the checker of the HIP path, never the product.
 * Each function names the reference routine (file:line under src/ of the reference) whose
 * arithmetic it reproduces, operation for operation, so that results are bit-identical.
 * Build: gcc -O2 -ffp-contract=off (no FMA contraction: the reference is built without it).
 */
#include "sayaka_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXLV 16
enum { FF = 0, FBC = 1, FC = 2, CF = 3 }; /* TowerBlockFaceType, SayakaTower.h:15-20 */
enum { NEUMANN = 0, DIRICHLET = 1, SIMPLEFILL = 2 };

typedef struct {
    int n;
    int* blocks;
    int begin[MAXLV + 1], end[MAXLV + 1];
    int *face_type, *face_neigh, *fc_off, *cf_neigh, *cf_off, *cf_box; /* indexed by block id */
} tower_t;

struct oc_ctx {
    int ndim, nb, nbz, nblk, minlevel, maxlevel;
    int nf;                 /* faces per block: 2*ndim */
    int nchild;             /* 2^ndim */
    int ncell;              /* nb*nb*nbz */
    int gx, gy, gz, ngrown; /* grown box extents */
    int kg;                 /* ghost layers in z: 1 in 3D, 0 in 2D */
    int fd[3][3];           /* face box extents per direction */
    int nface[3];
    int *level, *parent, *which_child, *child, *nbr, *node_type;
    double *dx, *ds, *dv;
    int *lvl_start, *lvl_blocks;
    tower_t tw[MAXLV + 1];
    double* buf;
    double* save[3];
    double* vf;
    double *af[3], *bcoef[3];
    int* bc_type;
    double* bc_val;
    double alpha, beta, tol_rel, tol_abs, bot_tol_rel, bot_tol_abs;
    int max_iter, bot_max_iter, precond, post_smooth;
    int bottom_calls;
    long long bottom_iters;
};

/* stencil and stagger offsets per face (SayakaMacOperator.cpp:1170-1188) */
static const int IOFF[6] = {-1, 1, 0, 0, 0, 0}, JOFF[6] = {0, 0, -1, 1, 0, 0}, KOFF[6] = {0, 0, 0, 0, -1, 1};
static const int IST[6] = {0, 1, 0, 0, 0, 0}, JST[6] = {0, 0, 0, 1, 0, 0}, KST[6] = {0, 0, 0, 0, 0, 1};

static inline int is_leaf(const oc_ctx* c, int b) { return c->node_type[b] == 1; }
static inline int gidx(const oc_ctx* c, int i, int j, int k) {
    return (i + 1) + c->gx * ((j + 1) + c->gy * (k + c->kg));
}
static inline int vidx(const oc_ctx* c, int i, int j, int k) { return i + c->nb * (j + c->nb * k); }
static inline int fidx(const oc_ctx* c, int d, int i, int j, int k) {
    return i + c->fd[d][0] * (j + c->fd[d][1] * k);
}
static inline double* G(const oc_ctx* c, int comp, int b) {
    return c->buf + ((size_t)comp * c->nblk + b) * c->ngrown;
}
/* IndexMapping::FineToCrse / CrseToFine (SayakaTreeData.h:34-44), per active direction */
static inline int f2c(int i, int off) { return (i + 64) / 2 - 32 + off; }
static inline int c2f(int ic, int s, int off) { return (ic - off) * 2 + s; }
static inline int side_of(int child, int d) { return (child >> d) & 1; }

/* ------------------------------------------------------------------------------ set-up */
static void build_tower(oc_ctx* c, int m) {
    /* MGLevelTower::update / registerBlock (SayakaTower.cpp:28-187) */
    tower_t* t = &c->tw[m];
    const int nb = c->nb, NF = c->nf;
    t->n = 0;
    for (int l = 1; l <= MAXLV; l++) t->begin[l] = t->end[l] = 0;
    int lvcount = c->lvl_start[m + 1] - c->lvl_start[m];
    if (lvcount == 0) return;
    for (int l = m; l >= 1; l--) {
        t->begin[l] = t->n;
        for (int q = c->lvl_start[l]; q < c->lvl_start[l + 1]; q++) {
            const int b = c->lvl_blocks[q];
            if (!(is_leaf(c, b) || l == m)) continue;
            t->blocks[t->n++] = b;
            for (int f = 0; f < NF; f++) {
                const int nn = c->nbr[b * 6 + f];
                int* ft = &t->face_type[b * 6 + f];
                int* fn = &t->face_neigh[b * 6 + f];
                if (nn <= -20) { *ft = FBC; *fn = nn; }
                else if (nn < 0) { *ft = FC; *fn = c->nbr[c->parent[b] * 6 + f]; }
                else if (l < m && !is_leaf(c, nn)) { *ft = CF; *fn = nn; }
                else { *ft = FF; *fn = nn; }
            }
            for (int f = 0; f < NF; f++) {
                const int d = f / 2, s = f % 2;
                const int ft = t->face_type[b * 6 + f];
                if (ft == FC) {
                    const int wc = c->which_child[b];
                    const int nc = wc ^ (1 << d); /* ChildIndex::neighbor across f */
                    for (int a = 0; a < 3; a++) t->fc_off[(b * 6 + f) * 3 + a] = a < c->ndim ? side_of(nc, a) * (nb / 2) : 0;
                } else if (ft == CF) {
                    int cnt = 0;
                    const int nn = t->face_neigh[b * 6 + f];
                    for (int ch = 0; ch < c->nchild; ch++) {
                        if (side_of(ch, d) != 1 - s) continue; /* hasFaceBoundary(opposite) */
                        int* off = &t->cf_off[((b * 6 + f) * 4 + cnt) * 3];
                        int* box = &t->cf_box[((b * 6 + f) * 4 + cnt) * 6];
                        for (int a = 0; a < 3; a++) {
                            const int ext = a == 2 ? c->nbz : nb;
                            off[a] = a < c->ndim ? side_of(ch, a) * (nb / 2) : 0;
                            if (a < c->ndim) {
                                const int mid = (0 + ext - 1) / 2;
                                box[a] = side_of(ch, a) ? mid + 1 : 0;
                                box[3 + a] = side_of(ch, a) ? ext - 1 : mid;
                            } else {
                                box[a] = 0;
                                box[3 + a] = ext - 1;
                            }
                        }
                        t->cf_neigh[(b * 6 + f) * 4 + cnt] = c->child[nn * 8 + ch];
                        cnt++;
                    }
                }
            }
        }
        t->end[l] = t->n;
    }
}

oc_ctx* oc_create(int ndim, int nb, int nblk, int minlevel, int maxlevel, const int* level,
                  const int* parent, const int* which_child, const int* child8, const int* nbr6,
                  const int* node_type, const double* length3) {
    if ((ndim != 2 && ndim != 3) || nb < 2 || maxlevel > MAXLV - 1) return NULL;
    oc_ctx* c = (oc_ctx*)calloc(1, sizeof(oc_ctx));
    c->ndim = ndim;
    c->nb = nb;
    c->nbz = ndim == 3 ? nb : 1;
    c->nblk = nblk;
    c->minlevel = minlevel;
    c->maxlevel = maxlevel;
    c->nf = 2 * ndim;
    c->nchild = 1 << ndim;
    c->ncell = nb * nb * c->nbz;
    c->kg = ndim == 3 ? 1 : 0;
    c->gx = nb + 2;
    c->gy = nb + 2;
    c->gz = c->nbz + 2 * c->kg;
    c->ngrown = c->gx * c->gy * c->gz;
    for (int d = 0; d < 3; d++) {
        c->fd[d][0] = nb;
        c->fd[d][1] = nb;
        c->fd[d][2] = c->nbz;
        if (d < ndim) c->fd[d][d] += 1;
        c->nface[d] = c->fd[d][0] * c->fd[d][1] * c->fd[d][2];
    }
    const size_t n = (size_t)nblk;
    c->level = malloc(n * sizeof(int));
    c->parent = malloc(n * sizeof(int));
    c->which_child = malloc(n * sizeof(int));
    c->child = malloc(n * 8 * sizeof(int));
    c->nbr = malloc(n * 6 * sizeof(int));
    c->node_type = malloc(n * sizeof(int));
    memcpy(c->level, level, n * sizeof(int));
    memcpy(c->parent, parent, n * sizeof(int));
    memcpy(c->which_child, which_child, n * sizeof(int));
    memcpy(c->child, child8, n * 8 * sizeof(int));
    memcpy(c->nbr, nbr6, n * 6 * sizeof(int));
    memcpy(c->node_type, node_type, n * sizeof(int));
    /* block metric (SayakaTree.h:277-301): dx = length/nb, dv = prod dx, ds = dv/dx */
    c->dx = malloc(n * 3 * sizeof(double));
    c->ds = malloc(n * 3 * sizeof(double));
    c->dv = malloc(n * sizeof(double));
    for (int b = 0; b < nblk; b++) {
        double dx[3], v = 1.0;
        for (int d = 0; d < 3; d++) dx[d] = length3[b * 3 + d];
        for (int d = 0; d < ndim; d++) dx[d] /= nb;
        for (int d = 0; d < ndim; d++) v *= dx[d];
        for (int d = 0; d < 3; d++) {
            c->dx[b * 3 + d] = dx[d];
            c->ds[b * 3 + d] = d < ndim ? v / dx[d] : 1.0;
        }
        c->dv[b] = v;
    }
    /* per-level block lists in ascending id (SayakaRegrid.cpp:962-976) */
    c->lvl_start = calloc(MAXLV + 2, sizeof(int));
    c->lvl_blocks = malloc((n + 1) * sizeof(int));
    for (int b = 0; b < nblk; b++) c->lvl_start[level[b] + 1]++;
    for (int l = 1; l <= MAXLV + 1; l++) c->lvl_start[l] += c->lvl_start[l - 1];
    {
        int* fill = calloc(MAXLV + 2, sizeof(int));
        for (int b = 0; b < nblk; b++) c->lvl_blocks[c->lvl_start[level[b]] + fill[level[b]]++] = b;
        free(fill);
    }
    for (int m = 1; m <= maxlevel; m++) {
        tower_t* t = &c->tw[m];
        t->blocks = malloc((n + 1) * sizeof(int));
        t->face_type = calloc(n * 6, sizeof(int));
        t->face_neigh = calloc(n * 6, sizeof(int));
        t->fc_off = calloc(n * 18, sizeof(int));
        t->cf_neigh = calloc(n * 24, sizeof(int));
        t->cf_off = calloc(n * 72, sizeof(int));
        t->cf_box = calloc(n * 144, sizeof(int));
    }
    c->buf = calloc((size_t)OC_NCOMP * n * c->ngrown, sizeof(double));
    c->vf = calloc(n * c->ncell, sizeof(double));
    for (int d = 0; d < 3; d++) {
        c->save[d] = calloc(n * c->nface[d], sizeof(double));
        c->af[d] = calloc(n * c->nface[d], sizeof(double));
        c->bcoef[d] = calloc(n * c->nface[d], sizeof(double));
    }
    c->bc_type = malloc(n * 6 * sizeof(int));
    c->bc_val = calloc(n * 6, sizeof(double));
    for (size_t q = 0; q < n * 6; q++) c->bc_type[q] = -1;
    /* MacSolver constructor defaults (SayakaMacSolver.h:56-87) */
    c->alpha = 0.0;
    c->beta = 1.0;
    c->max_iter = 2000;
    c->tol_rel = 1.0e-6;
    c->tol_abs = 0.0;
    c->post_smooth = 2;
    c->bot_tol_abs = 1.0e-6;
    c->bot_tol_rel = 1.0e-6;
    c->bot_max_iter = 2000;
    c->precond = 0;
    return c;
}

void oc_destroy(oc_ctx* c) {
    if (!c) return;
    free(c->level); free(c->parent); free(c->which_child); free(c->child); free(c->nbr);
    free(c->node_type); free(c->dx); free(c->ds); free(c->dv); free(c->lvl_start); free(c->lvl_blocks);
    for (int m = 1; m <= c->maxlevel; m++) {
        tower_t* t = &c->tw[m];
        free(t->blocks); free(t->face_type); free(t->face_neigh); free(t->fc_off);
        free(t->cf_neigh); free(t->cf_off); free(t->cf_box);
    }
    free(c->buf); free(c->vf);
    for (int d = 0; d < 3; d++) { free(c->save[d]); free(c->af[d]); free(c->bcoef[d]); }
    free(c->bc_type); free(c->bc_val);
    free(c);
}

int oc_tower_size(const oc_ctx* c, int m) { return c->tw[m].n; }

void oc_tower_export(const oc_ctx* c, int m, int* blocks, int* begin, int* end, int* face_type6,
                     int* face_neigh6, int* fc_off18, int* cf_neigh24, int* cf_off72, int* cf_box144) {
    const tower_t* t = &c->tw[m];
    for (int l = 1; l <= m; l++) { begin[l] = t->begin[l]; end[l] = t->end[l]; }
    for (int q = 0; q < t->n; q++) {
        const int b = t->blocks[q];
        blocks[q] = b;
        for (int f = 0; f < c->nf; f++) {
            const int ft = t->face_type[b * 6 + f];
            face_type6[q * 6 + f] = ft;
            face_neigh6[q * 6 + f] = t->face_neigh[b * 6 + f];
            for (int a = 0; a < 3; a++) fc_off18[(q * 6 + f) * 3 + a] = ft == FC ? t->fc_off[(b * 6 + f) * 3 + a] : 0;
            for (int k = 0; k < 4; k++) {
                const int ok = ft == CF && k < c->nchild / 2;
                cf_neigh24[(q * 6 + f) * 4 + k] = ok ? t->cf_neigh[(b * 6 + f) * 4 + k] : -1;
                for (int a = 0; a < 3; a++) cf_off72[((q * 6 + f) * 4 + k) * 3 + a] = ok ? t->cf_off[((b * 6 + f) * 4 + k) * 3 + a] : 0;
                for (int a = 0; a < 6; a++) cf_box144[((q * 6 + f) * 4 + k) * 6 + a] = ok ? t->cf_box[((b * 6 + f) * 4 + k) * 6 + a] : 0;
            }
        }
    }
}

/* -------------------------------------------------------------------- face-data helpers */
/* restrictFaceBlockData_Average over the face box of child `ch` on side `s` of direction d
 * (SayakaInterpPatch.cpp:627-691); writes the coarse block `pb`'s face field. */
static void restrict_face_box(const oc_ctx* c, double* fdata, int d, int pb, int fb, int ch,
                              const int lo[3], const int hi[3], int do_avg) {
    const int nb = c->nb;
    const int off[3] = {side_of(ch, 0) * (nb / 2), side_of(ch, 1) * (nb / 2), c->ndim == 3 ? side_of(ch, 2) * (nb / 2) : 0};
    const int t1[3] = {d == 0 ? 0 : 1, d == 1 ? 0 : 1, (d == 2 || c->ndim < 3) ? 0 : 1};
    const double scale = 1.0 / ((2 * 2 * (c->ndim == 3 ? 2 : 1)) / 2);
    double* cd = fdata + (size_t)pb * c->nface[d];
    const double* fdv = fdata + (size_t)fb * c->nface[d];
    for (int k = lo[2]; k <= hi[2]; k++)
        for (int j = lo[1]; j <= hi[1]; j++)
            for (int i = lo[0]; i <= hi[0]; i++) {
                double val = 0;
                for (int kind = -t1[2]; kind <= t1[2]; kind += 2) {
                    const int kf = c->ndim == 3 ? c2f(k, (kind + 1) / 2, off[2]) : k;
                    for (int jind = -t1[1]; jind <= t1[1]; jind += 2) {
                        const int jf = c2f(j, (jind + 1) / 2, off[1]);
                        for (int iind = -t1[0]; iind <= t1[0]; iind += 2) {
                            const int ifn = c2f(i, (iind + 1) / 2, off[0]);
                            val += fdv[fidx(c, d, ifn, jf, kf)];
                        }
                    }
                }
                cd[fidx(c, d, i, j, k)] = do_avg ? val * scale : val;
            }
}

/* ChildSubIndexRange of child `ch` (SayakaChildIndex.cpp:373-393) */
static void child_box(const oc_ctx* c, int ch, int lo[3], int hi[3]) {
    const int ext[3] = {c->nb, c->nb, c->nbz};
    for (int a = 0; a < 3; a++) {
        lo[a] = 0;
        hi[a] = ext[a] - 1;
        if (a < c->ndim) {
            const int mid = (ext[a] - 1) / 2;
            if (side_of(ch, a)) lo[a] = mid + 1;
            else hi[a] = mid;
        }
    }
}

/* MGLevelTower::syncSubLevelCrseFineFlux (SayakaTower.cpp:298-415) */
static void sync_cf(const oc_ctx* c, int m, int fl, int cl, int d, double* fdata, int do_avg) {
    const tower_t* t = &c->tw[m];
    if (t->n == 0 || t->end[fl] - t->begin[fl] <= 0 || t->end[cl] - t->begin[cl] <= 0) return;
    for (int q = t->begin[fl]; q < t->end[fl]; q++) {
        const int b = t->blocks[q];
        for (int s = 0; s <= 1; s++) {
            if (t->face_type[b * 6 + 2 * d + s] != FC) continue;
            const int ch = c->which_child[b];
            int lo[3], hi[3];
            child_box(c, ch, lo, hi);
            /* BoundaryFace: staggered in d, collapsed to side s */
            if (s == 0) hi[d] = lo[d];
            else { hi[d] += 1; lo[d] = hi[d]; }
            restrict_face_box(c, fdata, d, c->parent[b], b, ch, lo, hi, do_avg);
        }
    }
    const int nf = c->nface[d];
    for (int q = t->begin[cl]; q < t->end[cl]; q++) {
        const int b = t->blocks[q];
        for (int s = 0; s <= 1; s++) {
            if (t->face_type[b * 6 + 2 * d + s] != CF) continue;
            const int nn = c->nbr[b * 6 + 2 * d + s];
            double* dst = fdata + (size_t)b * nf;
            const double* src = fdata + (size_t)nn * nf;
            int lo[3] = {0, 0, 0}, hi[3] = {c->fd[d][0] - 1, c->fd[d][1] - 1, c->fd[d][2] - 1};
            int so = s == 0 ? 0 : hi[d];          /* own face slab */
            int sn = s == 0 ? hi[d] : 0;          /* neighbour's opposite slab */
            lo[d] = hi[d] = 0;
            for (int k = lo[2]; k <= hi[2]; k++)
                for (int j = lo[1]; j <= hi[1]; j++)
                    for (int i = lo[0]; i <= hi[0]; i++) {
                        int io[3] = {i, j, k}, in[3] = {i, j, k};
                        io[d] = so;
                        in[d] = sn;
                        dst[fidx(c, d, io[0], io[1], io[2])] = src[fidx(c, d, in[0], in[1], in[2])];
                    }
        }
    }
}

/* FillPatch::restrictLevel / restrictBlock for face data, average (SayakaFillPatch.cpp:401-447) */
static void restrict_face_level(const oc_ctx* c, double* fdata, int d, int cl) {
    if (cl == c->maxlevel) return;
    for (int q = c->lvl_start[cl]; q < c->lvl_start[cl + 1]; q++) {
        const int b = c->lvl_blocks[q];
        if (is_leaf(c, b)) continue;
        for (int ch = 0; ch < c->nchild; ch++) {
            int lo[3], hi[3];
            child_box(c, ch, lo, hi);
            hi[d] += 1; /* staggerInDir */
            restrict_face_box(c, fdata, d, b, c->child[b * 8 + ch], ch, lo, hi, 1);
        }
    }
}

/* cell average restriction of a valid-layout field (restrictCellBlockData_Average,
 * SayakaInterpPatch.cpp:398-444) over all non-leaf blocks of level cl */
static void restrict_cell_level_valid(const oc_ctx* c, double* data, int cl) {
    if (cl == c->maxlevel) return;
    const int nb = c->nb;
    const double wgt = 1.0 / c->nchild;
    for (int q = c->lvl_start[cl]; q < c->lvl_start[cl + 1]; q++) {
        const int b = c->lvl_blocks[q];
        if (is_leaf(c, b)) continue;
        for (int ch = 0; ch < c->nchild; ch++) {
            const int fb = c->child[b * 8 + ch];
            int lo[3], hi[3];
            child_box(c, ch, lo, hi);
            const int off[3] = {side_of(ch, 0) * (nb / 2), side_of(ch, 1) * (nb / 2), c->ndim == 3 ? side_of(ch, 2) * (nb / 2) : 0};
            for (int kc = lo[2]; kc <= hi[2]; kc++)
                for (int jc = lo[1]; jc <= hi[1]; jc++)
                    for (int ic = lo[0]; ic <= hi[0]; ic++) {
                        double sum = 0;
                        for (int kk = 0; kk <= (c->ndim == 3); kk++) {
                            const int kf = c->ndim == 3 ? c2f(kc, kk, off[2]) : kc;
                            for (int jj = 0; jj <= 1; jj++) {
                                const int jf = c2f(jc, jj, off[1]);
                                for (int ii = 0; ii <= 1; ii++) {
                                    const int ifn = c2f(ic, ii, off[0]);
                                    sum += data[(size_t)fb * c->ncell + vidx(c, ifn, jf, kf)];
                                }
                            }
                        }
                        data[(size_t)b * c->ncell + vidx(c, ic, jc, kc)] = sum * wgt;
                    }
        }
    }
}

/* ----------------------------------------------------------------------------- set-up */
void oc_prepare(oc_ctx* c) {
    /* registerMGLevels (SayakaMacOperator.cpp:62-78): towers maxlevel..1 */
    for (int m = c->maxlevel; m >= 1; m--) build_tower(c, m);
    /* setDefaultCoefs, setUnitIBVolumeFrac/AreaFrac (SayakaMacCoef.cpp:8-18, SayakaMacIB.cpp:16-32) */
    c->alpha = 0.0;
    c->beta = 1.0;
    for (int d = 0; d < c->ndim; d++)
        for (size_t q = 0; q < (size_t)c->nblk * c->nface[d]; q++) { c->bcoef[d][q] = 1.0; c->af[d][q] = 1.0; }
    for (size_t q = 0; q < (size_t)c->nblk * c->ncell; q++) c->vf[q] = 1.0;
}

static int finest(const oc_ctx* c) {
    int f = 0;
    for (int b = 0; b < c->nblk; b++) if (c->level[b] > f) f = c->level[b];
    return f;
}

/* MacSolver::setCoefB (SayakaMacCoef.cpp:319-367) */
void oc_set_coefb(oc_ctx* c, const double* const* b) {
    const int L = finest(c);
    const tower_t* t = &c->tw[L];
    for (int d = 0; d < c->ndim; d++) {
        const int nf = c->nface[d];
        for (int q = 0; q < t->n; q++) {
            const int blk = t->blocks[q];
            memcpy(c->bcoef[d] + (size_t)blk * nf, b[d] + (size_t)blk * nf, nf * sizeof(double));
        }
        for (int l = L; l >= 2; l--) sync_cf(c, L, l, l - 1, d, c->bcoef[d], 1);
        for (int l = L - 1; l >= 1; l--) restrict_face_level(c, c->bcoef[d], d, l);
    }
}

/* steps of MacSolver::calcIBFracByLS after the leaf values (SayakaMacIB.cpp:69-96) */
static void ib_finish(oc_ctx* c) {
    const int L = finest(c);
    for (int l = L; l >= 2; l--)
        for (int d = 0; d < c->ndim; d++) sync_cf(c, L, l, l - 1, d, c->af[d], 1);
    for (int l = L - 1; l >= 1; l--) restrict_cell_level_valid(c, c->vf, l);
    for (int d = 0; d < c->ndim; d++)
        for (int l = L - 1; l >= 1; l--) restrict_face_level(c, c->af[d], d, l);
}

void oc_set_ib_frac(oc_ctx* c, const double* vf, const double* const* af) {
    const int L = finest(c);
    const tower_t* t = &c->tw[L];
    for (int q = 0; q < t->n; q++) {
        const int b = t->blocks[q];
        memcpy(c->vf + (size_t)b * c->ncell, vf + (size_t)b * c->ncell, c->ncell * sizeof(double));
        for (int d = 0; d < c->ndim; d++)
            memcpy(c->af[d] + (size_t)b * c->nface[d], af[d] + (size_t)b * c->nface[d], c->nface[d] * sizeof(double));
    }
    ib_finish(c);
}

/* calcBlockIBVolumeFrac / calcBlockIBAreaFrac: solid where ls <= 0 (SayakaMacIB.cpp:99-154) */
void oc_set_ib_ls(oc_ctx* c, const double* ls, int ng) {
    const int L = finest(c);
    const tower_t* t = &c->tw[L];
    const int nb = c->nb, nbz = c->nbz, kg = c->ndim == 3 ? ng : 0;
    const int lx = nb + 2 * ng, ly = nb + 2 * ng, lz = nbz + 2 * kg;
#define LS(b, i, j, k) ls[(size_t)(b) * lx * ly * lz + ((i) + ng) + lx * (((j) + ng) + ly * ((k) + kg))]
    for (int q = 0; q < t->n; q++) {
        const int b = t->blocks[q];
        for (int k = 0; k < nbz; k++)
            for (int j = 0; j < nb; j++)
                for (int i = 0; i < nb; i++) c->vf[(size_t)b * c->ncell + vidx(c, i, j, k)] = LS(b, i, j, k) <= 0 ? 0.0 : 1.0;
        for (int d = 0; d < c->ndim; d++) {
            const int ii = d == 0, jj = d == 1, kk = d == 2;
            for (int k = 0; k < c->fd[d][2]; k++)
                for (int j = 0; j < c->fd[d][1]; j++)
                    for (int i = 0; i < c->fd[d][0]; i++)
                        c->af[d][(size_t)b * c->nface[d] + fidx(c, d, i, j, k)] =
                            (LS(b, i - ii, j - jj, k - kk) <= 0 || LS(b, i, j, k) <= 0) ? 0.0 : 1.0;
        }
    }
#undef LS
    ib_finish(c);
}

void oc_get_coefs(const oc_ctx* c, double* vf, double* const* b, double* const* af) {
    if (vf) memcpy(vf, c->vf, (size_t)c->nblk * c->ncell * sizeof(double));
    for (int d = 0; d < c->ndim; d++) {
        if (b && b[d]) memcpy(b[d], c->bcoef[d], (size_t)c->nblk * c->nface[d] * sizeof(double));
        if (af && af[d]) memcpy(af[d], c->af[d], (size_t)c->nblk * c->nface[d] * sizeof(double));
    }
}

/* MacSolver::setBC -> BlockBCRecord::setBlock (SayakaMacOperator.cpp:1462-1469,
 * SayakaBoundaryPatch.h:148-167) with one map for every surrounding position */
void oc_set_bc(oc_ctx* c, int n, const int* physbc, const int* mathbc, const double* val) {
    for (int b = 0; b < c->nblk; b++)
        for (int f = 0; f < 6; f++) {
            c->bc_type[b * 6 + f] = -1;
            c->bc_val[b * 6 + f] = 0.0;
            if (f >= c->nf) continue;
            const int p = c->nbr[b * 6 + f];
            if (p > -20) continue;
            for (int q = 0; q < n; q++)
                if (physbc[q] == p) { c->bc_type[b * 6 + f] = mathbc[q]; c->bc_val[b * 6 + f] = val[q]; }
        }
}

/* MacSolver::fixBC / fixBlockBC (SayakaMacOperator.cpp:1470-1594) */
void oc_fix_bc(const oc_ctx* c, const double* phi, int ng, double* rhs) {
    const int L = finest(c);
    const tower_t* t = &c->tw[L];
    const int nb = c->nb, nbz = c->nbz, kg = c->ndim == 3 ? ng : 0;
    const int lx = nb + 2 * ng, ly = nb + 2 * ng, lz = nbz + 2 * kg;
#define PHI(b, i, j, k) phi[(size_t)(b) * lx * ly * lz + ((i) + ng) + lx * (((j) + ng) + ly * ((k) + kg))]
    for (int q = 0; q < t->n; q++) {
        const int b = t->blocks[q];
        for (int f = 0; f < c->nf; f++) {
            if (t->face_type[b * 6 + f] != FBC) continue;
            const int bct = c->bc_type[b * 6 + f];
            const int d = f / 2, s = f % 2;
            int lo[3] = {0, 0, 0}, hi[3] = {nb - 1, nb - 1, nbz - 1};
            lo[d] = hi[d] = s == 0 ? 0 : hi[d];
            for (int k = lo[2]; k <= hi[2]; k++)
                for (int j = lo[1]; j <= hi[1]; j++)
                    for (int i = lo[0]; i <= hi[0]; i++) {
                        const double phi0 = PHI(b, i, j, k);
                        const double phi1 = PHI(b, i + IOFF[f], j + JOFF[f], k + KOFF[f]);
                        double coef = c->beta * c->ds[b * 3 + d] / c->dx[b * 3 + d];
                        const int fi = fidx(c, d, i + IST[f], j + JST[f], k + KST[f]);
                        coef *= c->bcoef[d][(size_t)b * c->nface[d] + fi];
                        coef *= c->af[d][(size_t)b * c->nface[d] + fi];
                        double corr = 0;
                        if (bct == DIRICHLET) {
                            const double phibc = 0.5 * (phi0 + phi1);
                            corr = phibc * 2.0 * coef;
                        } else if (bct == SIMPLEFILL) {
                            corr = phi1 * 2.0 * coef;
                        }
                        const double vfv = c->vf[(size_t)b * c->ncell + vidx(c, i, j, k)];
                        if (vfv > 0) {
                            corr /= (c->dv[b] * vfv);
                            rhs[(size_t)b * c->ncell + vidx(c, i, j, k)] += corr;
                        }
                    }
        }
    }
#undef PHI
}

void oc_set_params(oc_ctx* c, double tol_rel, double tol_abs, int max_iter, int precond) {
    c->tol_rel = tol_rel;
    c->tol_abs = tol_abs;
    c->max_iter = max_iter;
    c->precond = precond;
}

static void valid_in(oc_ctx* c, int comp, const double* v) {
    for (int b = 0; b < c->nblk; b++) {
        double* g = G(c, comp, b);
        for (int k = 0; k < c->nbz; k++)
            for (int j = 0; j < c->nb; j++)
                for (int i = 0; i < c->nb; i++) g[gidx(c, i, j, k)] = v[(size_t)b * c->ncell + vidx(c, i, j, k)];
    }
}
void oc_set_rhs(oc_ctx* c, const double* rhs) { valid_in(c, OC_RHS, rhs); }
void oc_set_init_sol(oc_ctx* c, const double* sol) { valid_in(c, OC_SOL, sol); }
void oc_set_zero_sol(oc_ctx* c) {
    for (int b = 0; b < c->nblk; b++) {
        double* g = G(c, OC_SOL, b);
        for (int k = 0; k < c->nbz; k++)
            for (int j = 0; j < c->nb; j++)
                for (int i = 0; i < c->nb; i++) g[gidx(c, i, j, k)] = 0.0;
    }
}
void oc_get_sol(const oc_ctx* c, double* out) {
    for (int b = 0; b < c->nblk; b++) {
        const double* g = G(c, OC_SOL, b);
        for (int k = 0; k < c->nbz; k++)
            for (int j = 0; j < c->nb; j++)
                for (int i = 0; i < c->nb; i++) out[(size_t)b * c->ncell + vidx(c, i, j, k)] = g[gidx(c, i, j, k)];
    }
}
void oc_buf_set(oc_ctx* c, int comp, const double* g) {
    memcpy(G(c, comp, 0), g, (size_t)c->nblk * c->ngrown * sizeof(double));
}
void oc_buf_get(const oc_ctx* c, int comp, double* g) {
    memcpy(g, G(c, comp, 0), (size_t)c->nblk * c->ngrown * sizeof(double));
}
void oc_save_get(const oc_ctx* c, int d, double* out) {
    memcpy(out, c->save[d], (size_t)c->nblk * c->nface[d] * sizeof(double));
}

/* ------------------------------------------------------------------------- ghost fill */
/* same-level face copy (fillBlockBndryAtSameLevel, SayakaFillPatch.cpp:92-190), faces only */
static void fill_same_level(const oc_ctx* c, int comp, int b) {
    const int nb = c->nb, nbz = c->nbz;
    double* dst = G(c, comp, b);
    for (int f = 0; f < c->nf; f++) {
        const int nn = c->nbr[b * 6 + f];
        if (nn < 0) continue;
        const double* src = G(c, comp, nn);
        const int d = f / 2, s = f % 2;
        int lo[3] = {0, 0, 0}, hi[3] = {nb - 1, nb - 1, nbz - 1};
        const int ext = d == 2 ? nbz : nb;
        lo[d] = hi[d] = s == 0 ? -1 : ext;
        for (int k = lo[2]; k <= hi[2]; k++)
            for (int j = lo[1]; j <= hi[1]; j++)
                for (int i = lo[0]; i <= hi[0]; i++) {
                    int p[3] = {i, j, k};
                    p[d] += s == 0 ? ext : -ext;
                    dst[gidx(c, i, j, k)] = src[gidx(c, p[0], p[1], p[2])];
                }
    }
}

/* injection of face ghosts from the parent where the face neighbour is coarse
 * (fillBlockBndryFromCoarseLevel + prolongCellBlockData_Injection,
 *  SayakaFillPatch.cpp:192-299, SayakaInterpPatch.cpp:157-189) */
static void fill_from_coarse(const oc_ctx* c, int comp, int b) {
    const int nb = c->nb, nbz = c->nbz;
    double* dst = G(c, comp, b);
    const int p = c->parent[b];
    const int wc = c->which_child[b];
    const int off[3] = {side_of(wc, 0) * (nb / 2), side_of(wc, 1) * (nb / 2), c->ndim == 3 ? side_of(wc, 2) * (nb / 2) : 0};
    for (int f = 0; f < c->nf; f++) {
        const int nn = c->nbr[b * 6 + f];
        if (!(nn < 0 && nn > -20)) continue;
        const double* src = G(c, comp, p);
        const int d = f / 2, s = f % 2;
        int lo[3] = {0, 0, 0}, hi[3] = {nb - 1, nb - 1, nbz - 1};
        const int ext = d == 2 ? nbz : nb;
        lo[d] = hi[d] = s == 0 ? -1 : ext;
        for (int k = lo[2]; k <= hi[2]; k++) {
            const int kc = c->ndim == 3 ? f2c(k, off[2]) : k;
            for (int j = lo[1]; j <= hi[1]; j++) {
                const int jc = f2c(j, off[1]);
                for (int i = lo[0]; i <= hi[0]; i++) dst[gidx(c, i, j, k)] = src[gidx(c, f2c(i, off[0]), jc, kc)];
            }
        }
    }
}

/* MacSolver::mg_fillbndry_level (SayakaMacUtil.cpp:210-251) */
void oc_fill(oc_ctx* c, int m, int comp) {
    const tower_t* t = &c->tw[m];
    if (c->lvl_start[m + 1] - c->lvl_start[m] == 0) return;
    for (int l = 1; l <= m; l++) {
        if (l < m && t->end[l + 1] - t->begin[l + 1] <= 0) continue;
        for (int q = c->lvl_start[l]; q < c->lvl_start[l + 1]; q++) {
            const int b = c->lvl_blocks[q];
            fill_same_level(c, comp, b);
            fill_from_coarse(c, comp, b);
        }
    }
}

static void valid_in(oc_ctx* c, int comp, const double* v);

/* ------------------------------------------------------------------------ face flux */
/* MacSolver::mg_flux_level, ghost-fill path (SayakaMacOperator.cpp:161-203), with
 * mg_flux_block_usefill (SayakaMacOperator.cpp:929-1025): flux = -(phi(i) - phi(i-1)) / dh on
 * the faces of direction d of every block of tower m, sub-level by sub-level from m down to 1;
 * wall faces 0; the fine side of a coarse/fine face divides by 1.5 dh and reads the injected
 * coarse ghost; the coarse side is left to syncSubLevelCrseFineFlux(FLUX_SYNC_AVG) after each
 * finer sub-level.  phi: valid values [blk][cell]; out: [blk][face of d], zeroed first (the
 * reference zeroes the level, SayakaMacOperator.cpp:176, and the sync also writes the
 * restricted fine fluxes into the non-leaf parents' faces).  The oracle's OC_TMP component
 * carries phi with its ghosts and is restored afterwards. */
void oc_flux(oc_ctx* c, int m, int d, const double* phi, double* out) {
    const tower_t* t = &c->tw[m];
    const size_t gsz = (size_t)c->nblk * c->ngrown;
    double* saved = (double*)malloc(gsz * sizeof(double));
    memcpy(saved, G(c, OC_TMP, 0), gsz * sizeof(double));
    memset(G(c, OC_TMP, 0), 0, gsz * sizeof(double));
    valid_in(c, OC_TMP, phi);
    const int nf = c->nface[d];
    memset(out, 0, (size_t)c->nblk * nf * sizeof(double));
    if (c->lvl_start[m + 1] - c->lvl_start[m] > 0) {
        oc_fill(c, m, OC_TMP);
        const int ii = d == 0, jj = d == 1, kk = d == 2;
        const int ext = d == 2 ? c->nbz : c->nb;
        for (int l = m; l >= 1; l--) {
            for (int q = t->begin[l]; q < t->end[l]; q++) {
                const int b = t->blocks[q];
                const double dh = c->dx[b * 3 + d];
                const double dh_fine_crse = dh * 1.5;
                const double* ph = G(c, OC_TMP, b);
                double* fo = out + (size_t)b * nf;
                for (int k = 0; k < c->fd[d][2]; k++)
                    for (int j = 0; j < c->fd[d][1]; j++)
                        for (int i = 0; i < c->fd[d][0]; i++) {
                            const int ijk = d == 0 ? i : (d == 1 ? j : k);
                            const double dphi = ph[gidx(c, i, j, k)] - ph[gidx(c, i - ii, j - jj, k - kk)];
                            const int side = ijk == 0 ? 0 : (ijk == ext ? 1 : -1);
                            double v;
                            if (side < 0) {
                                v = -dphi / dh;
                            } else {
                                const int ft = t->face_type[b * 6 + 2 * d + side];
                                if (ft == FBC) v = 0;
                                else if (ft == FF) v = -dphi / dh;
                                else if (ft == FC) v = -dphi / dh_fine_crse;
                                else continue; /* CF: filled by the sync below */
                            }
                            fo[fidx(c, d, i, j, k)] = v;
                        }
            }
            if (l > 1) sync_cf(c, m, l, l - 1, d, out, 1);
        }
    }
    memcpy(G(c, OC_TMP, 0), saved, gsz * sizeof(double));
    free(saved);
}

/* ------------------------------------------------------------------- operator, smoother */
/* cell_eval_coef + IB/B scaling (SayakaMacOperator.cpp:1115-1167, 1256-1266, 1395-1405) */
static void eval_coef(const oc_ctx* c, const tower_t* t, int b, int i, int j, int k, double off[6],
                      int bcf[6], int fcf[6], int cff[6]) {
    const int nb = c->nb;
    const int range[6] = {0, nb - 1, 0, nb - 1, 0, c->nbz - 1};
    const double beta = c->beta;
    for (int f = 0; f < c->nf; f++) {
        const int d = f / 2;
        const int ijk = d == 0 ? i : (d == 1 ? j : k);
        const double dsd = c->ds[b * 3 + d], dxd = c->dx[b * 3 + d];
        bcf[f] = fcf[f] = cff[f] = 0;
        if (ijk == range[f]) {
            const int ft = t->face_type[b * 6 + f];
            if (ft == FF) off[f] = beta * dsd / dxd;
            else if (ft == FBC) {
                if (c->bc_type[b * 6 + f] == NEUMANN) { off[f] = 0.0; bcf[f] = 1; }
                else { off[f] = beta * dsd / dxd; bcf[f] = 2; }
            } else if (ft == FC) { off[f] = beta * dsd / (dxd * 1.5); fcf[f] = 1; }
            else { off[f] = beta * dsd / (dxd * 0.75); cff[f] = 1; }
        } else {
            off[f] = beta * dsd / dxd;
        }
        const int fi = fidx(c, d, i + IST[f], j + JST[f], k + KST[f]);
        off[f] *= c->bcoef[d][(size_t)b * c->nface[d] + fi];
        off[f] *= c->af[d][(size_t)b * c->nface[d] + fi];
    }
}

/* mg_apply_block_usefill (SayakaMacOperator.cpp:1325-1459) */
static void apply_block(oc_ctx* c, const tower_t* t, int b, int outc, int inc) {
    const double* phi = G(c, inc, b);
    double* out = G(c, outc, b);
    const double dv = c->dv[b];
    for (int k = 0; k < c->nbz; k++)
        for (int j = 0; j < c->nb; j++)
            for (int i = 0; i < c->nb; i++) {
                double diag = c->alpha * dv;
                double off[6];
                int bcf[6], fcf[6], cff[6];
                eval_coef(c, t, b, i, j, k, off, bcf, fcf, cff);
                double offdiag = 0;
                const double phic = phi[gidx(c, i, j, k)];
                for (int f = 0; f < c->nf; f++) {
                    double offval = 0;
                    if (cff[f] == 0) {
                        if (bcf[f] == 1) { offval = 0; off[f] = 0; }
                        else if (bcf[f] == 2) offval = -off[f] * phic;
                        else offval = off[f] * phi[gidx(c, i + IOFF[f], j + JOFF[f], k + KOFF[f])];
                    } else {
                        const int d = f / 2;
                        offval = c->save[d][(size_t)b * c->nface[d] + fidx(c, d, i + IST[f], j + JST[f], k + KST[f])];
                    }
                    diag += off[f];
                    offdiag += offval;
                }
                const double Lphi = (phic * diag - offdiag) / dv;
                const double vfv = c->vf[(size_t)b * c->ncell + vidx(c, i, j, k)];
                out[gidx(c, i, j, k)] = vfv > 0 ? Lphi / vfv : 0;
                for (int f = 0; f < c->nf; f++)
                    if (fcf[f]) {
                        const int d = f / 2;
                        c->save[d][(size_t)b * c->nface[d] + fidx(c, d, i + IST[f], j + JST[f], k + KST[f])] = off[f] * phic;
                    }
            }
}

/* MacSolver::mg_apply_level (SayakaMacOperator.cpp:84-137) */
void oc_apply(oc_ctx* c, int m, int outc, int inc) {
    if (c->lvl_start[m + 1] - c->lvl_start[m] == 0) return;
    const tower_t* t = &c->tw[m];
    oc_fill(c, m, inc);
    for (int l = m; l >= 1; l--) {
        for (int q = t->begin[l]; q < t->end[l]; q++) apply_block(c, t, t->blocks[q], outc, inc);
        if (l > 1)
            for (int d = 0; d < c->ndim; d++) sync_cf(c, m, l, l - 1, d, c->save[d], 0);
    }
}

/* MacSolver::mg_residual_level (SayakaMacUtil.cpp:9-49) */
void oc_residual(oc_ctx* c, int m, int rc, int xc, int rhsc) {
    oc_apply(c, m, rc, xc);
    const tower_t* t = &c->tw[m];
    for (int q = 0; q < t->n; q++) {
        const int b = t->blocks[q];
        double* r = G(c, rc, b);
        const double* rhs = G(c, rhsc, b);
        for (int k = 0; k < c->nbz; k++)
            for (int j = 0; j < c->nb; j++)
                for (int i = 0; i < c->nb; i++) {
                    const int g = gidx(c, i, j, k);
                    const double Lphi = r[g];
                    r[g] = c->vf[(size_t)b * c->ncell + vidx(c, i, j, k)] > 0 ? rhs[g] - Lphi : 0;
                }
    }
}

/* mg_smooth_block_rbcycle_usefill (SayakaMacOperator.cpp:1191-1323) */
static void smooth_block(oc_ctx* c, const tower_t* t, int b, int solc, int rhsc, int rb) {
    double* phi = G(c, solc, b);
    const double* rhs = G(c, rhsc, b);
    const double dv = c->dv[b];
    for (int k = 0; k < c->nbz; k++)
        for (int j = 0; j < c->nb; j++)
            for (int i = 0; i < c->nb; i++) {
                if ((i + j + k) % 2 != rb) continue;
                double diag = c->alpha * dv;
                double off[6];
                int bcf[6], fcf[6], cff[6];
                eval_coef(c, t, b, i, j, k, off, bcf, fcf, cff);
                double offdiag = 0;
                for (int f = 0; f < c->nf; f++) {
                    double offval = 0;
                    if (cff[f] == 0) {
                        if (bcf[f] == 1) { offval = 0; off[f] = 0; }
                        else if (bcf[f] == 2) { offval = 0; off[f] *= 2.0; }
                        else offval = off[f] * phi[gidx(c, i + IOFF[f], j + JOFF[f], k + KOFF[f])];
                    } else {
                        const int d = f / 2;
                        offval = c->save[d][(size_t)b * c->nface[d] + fidx(c, d, i + IST[f], j + JST[f], k + KST[f])];
                    }
                    diag += off[f];
                    offdiag += offval;
                }
                const double rhsval = rhs[gidx(c, i, j, k)];
                const double vfv = c->vf[(size_t)b * c->ncell + vidx(c, i, j, k)];
                const double unew = vfv > 0 ? (rhsval * dv * vfv + offdiag) / diag : 0;
                phi[gidx(c, i, j, k)] = unew;
                for (int f = 0; f < c->nf; f++)
                    if (fcf[f]) {
                        const int d = f / 2;
                        c->save[d][(size_t)b * c->nface[d] + fidx(c, d, i + IST[f], j + JST[f], k + KST[f])] = off[f] * unew;
                    }
            }
}

/* mg_smooth_level_usefill (SayakaMacOperator.cpp:1028-1112) */
void oc_smooth(oc_ctx* c, int m, int solc, int rhsc) {
    if (c->lvl_start[m + 1] - c->lvl_start[m] == 0) return;
    const tower_t* t = &c->tw[m];
    oc_fill(c, m, solc);
    for (int d = 0; d < c->ndim; d++) memset(c->save[d], 0, (size_t)c->nblk * c->nface[d] * sizeof(double));
    for (int l = m; l >= 1; l--) {
        if (t->begin[l] >= t->end[l]) break;
        for (int rb = 0; rb < 2; rb++) {
            for (int q = c->lvl_start[l]; q < c->lvl_start[l + 1]; q++) fill_same_level(c, solc, c->lvl_blocks[q]);
            for (int q = t->begin[l]; q < t->end[l]; q++) smooth_block(c, t, t->blocks[q], solc, rhsc, rb);
        }
        if (l > 1)
            for (int d = 0; d < c->ndim; d++) sync_cf(c, m, l, l - 1, d, c->save[d], 0);
    }
}

/* --------------------------------------------------------------- restrict / prolong */
/* MacSolver::mg_restict_wgt (SayakaMacIB.cpp:157-232) */
void oc_restrict(oc_ctx* c, int cl, int comp) {
    if (cl == c->maxlevel) return;
    if (c->lvl_start[cl + 1] - c->lvl_start[cl] == 0) return;
    const int fl = cl + 1;
    if (c->lvl_start[fl + 1] - c->lvl_start[fl] == 0) return;
    const int nb = c->nb, nbz = c->nbz;
    for (int q = c->lvl_start[fl]; q < c->lvl_start[fl + 1]; q++) {
        const int b = c->lvl_blocks[q];
        double* phi = G(c, comp, b);
        double* tmp = G(c, OC_TMP, b);
        const double* vf = c->vf + (size_t)b * c->ncell;
        for (int k = 0; k < nbz; k++)
            for (int j = 0; j < nb; j++)
                for (int i = 0; i < nb; i++) {
                    const int g = gidx(c, i, j, k);
                    tmp[g] = phi[g];
                    phi[g] *= vf[vidx(c, i, j, k)];
                }
    }
    const double wgt = 1.0 / c->nchild;
    for (int q = c->lvl_start[cl]; q < c->lvl_start[cl + 1]; q++) {
        const int b = c->lvl_blocks[q];
        if (is_leaf(c, b)) continue;
        double* cd = G(c, comp, b);
        for (int ch = 0; ch < c->nchild; ch++) {
            const double* fdv = G(c, comp, c->child[b * 8 + ch]);
            int lo[3], hi[3];
            child_box(c, ch, lo, hi);
            const int off[3] = {side_of(ch, 0) * (nb / 2), side_of(ch, 1) * (nb / 2), c->ndim == 3 ? side_of(ch, 2) * (nb / 2) : 0};
            for (int kc = lo[2]; kc <= hi[2]; kc++)
                for (int jc = lo[1]; jc <= hi[1]; jc++)
                    for (int ic = lo[0]; ic <= hi[0]; ic++) {
                        double sum = 0;
                        for (int kk = 0; kk <= (c->ndim == 3); kk++) {
                            const int kf = c->ndim == 3 ? c2f(kc, kk, off[2]) : kc;
                            for (int jj = 0; jj <= 1; jj++) {
                                const int jf = c2f(jc, jj, off[1]);
                                for (int ii = 0; ii <= 1; ii++) sum += fdv[gidx(c, c2f(ic, ii, off[0]), jf, kf)];
                            }
                        }
                        cd[gidx(c, ic, jc, kc)] = sum * wgt;
                    }
        }
        const double* vf = c->vf + (size_t)b * c->ncell;
        for (int k = 0; k < nbz; k++)
            for (int j = 0; j < nb; j++)
                for (int i = 0; i < nb; i++) {
                    const int g = gidx(c, i, j, k);
                    const double v = vf[vidx(c, i, j, k)];
                    if (v > 0) cd[g] /= v;
                    else cd[g] = 0;
                }
    }
    for (int q = c->lvl_start[fl]; q < c->lvl_start[fl + 1]; q++) {
        const int b = c->lvl_blocks[q];
        double* phi = G(c, comp, b);
        const double* tmp = G(c, OC_TMP, b);
        for (int k = 0; k < nbz; k++)
            for (int j = 0; j < nb; j++)
                for (int i = 0; i < nb; i++) phi[gidx(c, i, j, k)] = tmp[gidx(c, i, j, k)];
    }
}

/* MacSolver::mg_prolong_wgt (SayakaMacIB.cpp:234-283) */
void oc_prolong(oc_ctx* c, int fl, int comp) {
    if (fl == 1) return;
    if (c->lvl_start[fl + 1] - c->lvl_start[fl] == 0) return;
    const int nb = c->nb, nbz = c->nbz;
    for (int q = c->lvl_start[fl]; q < c->lvl_start[fl + 1]; q++) {
        const int b = c->lvl_blocks[q];
        const int p = c->parent[b];
        if (p < 0) continue;
        const int wc = c->which_child[b];
        const int off[3] = {side_of(wc, 0) * (nb / 2), side_of(wc, 1) * (nb / 2), c->ndim == 3 ? side_of(wc, 2) * (nb / 2) : 0};
        double* fdv = G(c, comp, b);
        const double* cd = G(c, comp, p);
        for (int k = 0; k < nbz; k++) {
            const int kc = c->ndim == 3 ? f2c(k, off[2]) : k;
            for (int j = 0; j < nb; j++) {
                const int jc = f2c(j, off[1]);
                for (int i = 0; i < nb; i++) fdv[gidx(c, i, j, k)] = cd[gidx(c, f2c(i, off[0]), jc, kc)];
            }
        }
    }
    for (int q = c->lvl_start[fl]; q < c->lvl_start[fl + 1]; q++) {
        const int b = c->lvl_blocks[q];
        double* fdv = G(c, comp, b);
        const double* vf = c->vf + (size_t)b * c->ncell;
        for (int k = 0; k < nbz; k++)
            for (int j = 0; j < nb; j++)
                for (int i = 0; i < nb; i++)
                    if (!(vf[vidx(c, i, j, k)] > 0)) fdv[gidx(c, i, j, k)] = 0;
    }
}

/* ------------------------------------------------------------------------------ BLAS-1 */
/* mg_norm_level -> ReduceValue(MAX_ABS) (SayakaMacUtil.cpp:53-76, SayakaBlockData.cpp:264-308) */
double oc_norm(const oc_ctx* c, int m, int a) {
    const tower_t* t = &c->tw[m];
    double norm = 0;
    for (int q = 0; q < t->n; q++) {
        const double* x = G(c, a, t->blocks[q]);
        double r = fabs(x[gidx(c, 0, 0, 0)]);
        for (int k = 0; k < c->nbz; k++)
            for (int j = 0; j < c->nb; j++)
                for (int i = 0; i < c->nb; i++) {
                    const double v = fabs(x[gidx(c, i, j, k)]);
                    r = r < v ? v : r; /* std::max(result, abs(val)) */
                }
        norm = norm < r ? r : norm;
    }
    return norm;
}

/* mg_dotprod_level -> DotProd (SayakaMacUtil.cpp:79-106, SayakaBlockData.cpp:185-208) */
double oc_dot(const oc_ctx* c, int m, int a, int b) {
    const tower_t* t = &c->tw[m];
    double dot = 0;
    for (int q = 0; q < t->n; q++) {
        const double* x = G(c, a, t->blocks[q]);
        const double* y = G(c, b, t->blocks[q]);
        double s = 0;
        for (int k = 0; k < c->nbz; k++)
            for (int j = 0; j < c->nb; j++)
                for (int i = 0; i < c->nb; i++) {
                    const int g = gidx(c, i, j, k);
                    s += x[g] * y[g];
                }
        dot += s;
    }
    return dot;
}

/* mg_saxby_level -> Calc_axpby (SayakaMacUtil.cpp:187-207, SayakaBlockData.cpp:212-261) */
static void saxby(oc_ctx* c, int m, int out, double a, int x, double b, int y) {
    const tower_t* t = &c->tw[m];
    for (int q = 0; q < t->n; q++) {
        const int blk = t->blocks[q];
        double* o = G(c, out, blk);
        const double* xv = G(c, x, blk);
        const double* yv = G(c, y, blk);
        for (int k = 0; k < c->nbz; k++)
            for (int j = 0; j < c->nb; j++)
                for (int i = 0; i < c->nb; i++) {
                    const int g = gidx(c, i, j, k);
                    const double v = a * xv[g] + b * yv[g];
                    o[g] = v;
                }
    }
}

/* mg_copy_level with ngrow 0 (SayakaMacUtil.cpp:175-185) */
static void copy_valid(oc_ctx* c, int m, int dst, int src) {
    const tower_t* t = &c->tw[m];
    for (int q = 0; q < t->n; q++) {
        const int b = t->blocks[q];
        double* o = G(c, dst, b);
        const double* s = G(c, src, b);
        for (int k = 0; k < c->nbz; k++)
            for (int j = 0; j < c->nb; j++)
                for (int i = 0; i < c->nb; i++) o[gidx(c, i, j, k)] = s[gidx(c, i, j, k)];
    }
}

/* mg_zero_level: whole grown block (SayakaMacUtil.cpp:136-150) */
static void zero_grown(oc_ctx* c, int m, int comp) {
    if (c->lvl_start[m + 1] - c->lvl_start[m] == 0) return;
    const tower_t* t = &c->tw[m];
    for (int q = 0; q < t->n; q++) memset(G(c, comp, t->blocks[q]), 0, c->ngrown * sizeof(double));
}

/* mg_correct_level -> AddEqual (SayakaMacUtil.cpp:109-133, SayakaBlockData.cpp:118-140) */
static void correct(oc_ctx* c, int m, int sol, int corr) {
    const tower_t* t = &c->tw[m];
    for (int q = 0; q < t->n; q++) {
        const int b = t->blocks[q];
        double* s = G(c, sol, b);
        const double* e = G(c, corr, b);
        for (int k = 0; k < c->nbz; k++)
            for (int j = 0; j < c->nb; j++)
                for (int i = 0; i < c->nb; i++) s[gidx(c, i, j, k)] += e[gidx(c, i, j, k)];
    }
}

/* ------------------------------------------------------------------------ solvers */
static int cg_level(oc_ctx* c, int m, int solc, int rhsc, double tol_rel, double tol_abs, int max_iter,
                    int precond, int* iters, int* converged, double* hist, int hist_cap, int* nhist);

/* MacSolver::mg_solve_bottom (SayakaMacSolver.cpp:164-242) */
static int bottom(oc_ctx* c) {
    zero_grown(c, 1, OC_CORR);
    int it = 0, conv = 0;
    int ret = cg_level(c, 1, OC_CORR, OC_RESID, c->bot_tol_rel, c->bot_tol_abs, c->bot_max_iter, 0, &it, &conv,
                       NULL, 0, NULL);
    c->bottom_calls++;
    c->bottom_iters += it;
    return ret == 0 ? 0 : 1;
}

/* MacSolver::mg_relax_vcycle with max_iter = fixed_iter = 1 (SayakaMacSolver.cpp:245-374) */
static int vcycle(oc_ctx* c, int L, int solc, int rhsc) {
    copy_valid(c, L, OC_SOL, solc);
    oc_residual(c, L, OC_RESID, OC_SOL, rhsc);
    (void)oc_norm(c, L, OC_RESID);
    zero_grown(c, L, OC_CORR);
    for (int l = L; l >= 2; l--) oc_restrict(c, l - 1, OC_RESID);
    int ret = 0;
    if (bottom(c) != 0) {
        ret = 1;
    } else {
        for (int l = 2; l <= L; l++) {
            oc_prolong(c, l, OC_CORR);
            for (int s = 1; s <= c->post_smooth; s++) oc_smooth(c, l, OC_CORR, OC_RESID);
        }
        correct(c, L, OC_SOL, OC_CORR);
        oc_residual(c, L, OC_RESID, OC_SOL, OC_RHS);
        (void)oc_norm(c, L, OC_RESID);
    }
    copy_valid(c, L, solc, OC_SOL);
    return ret;
}

/* MacSolver::mg_precond_level (SayakaMacSolver.cpp:924-949) */
int oc_precond(oc_ctx* c, int m, int solc, int rhsc) {
    zero_grown(c, m, solc);
    return vcycle(c, m, solc, rhsc);
}

/* MacSolver::cg_solve_level (SayakaMacSolver.cpp:683-920) */
static int cg_level(oc_ctx* c, int m, int solc, int rhsc, double tol_rel, double tol_abs, int max_iter,
                    int precond, int* iters, int* converged, double* hist, int hist_cap, int* nhist) {
    int stat = 0, nh = 0;
    copy_valid(c, m, OC_X0, solc);
    oc_residual(c, m, OC_R, OC_X0, rhsc);
    copy_valid(c, m, OC_RHAT, OC_R);
    zero_grown(c, m, OC_X);
    const double rnorm0 = oc_norm(c, m, OC_R);
    if (iters) *iters = 0;
    if (converged) *converged = 0;
    if (nhist) *nhist = 0;
    if (rnorm0 == 0) return stat;
    double rho1 = 0, alpha = 0, omega = 0, rnorm = 0;
    int conv = 0, iter;
    if (precond && m == 1) precond = 0;
    for (iter = 1; iter <= max_iter; iter++) {
        if (iters) *iters = iter;
        const double rho = oc_dot(c, m, OC_RHAT, OC_R);
        if (rho == 0) { stat = 1; break; }
        if (iter == 1) {
            copy_valid(c, m, OC_P, OC_R);
        } else {
            const double beta = (rho / rho1) * (alpha / omega);
            saxby(c, m, OC_P, 1.0, OC_P, -omega, OC_V);
            saxby(c, m, OC_P, 1.0, OC_R, beta, OC_P);
        }
        if (!precond) copy_valid(c, m, OC_PHAT, OC_P);
        else oc_precond(c, m, OC_PHAT, OC_P);
        oc_apply(c, m, OC_V, OC_PHAT);
        alpha = oc_dot(c, m, OC_RHAT, OC_V);
        if (alpha == 0) { stat = 1; break; }
        alpha = rho / alpha;
        saxby(c, m, OC_S, 1.0, OC_R, -alpha, OC_V);
        saxby(c, m, OC_X, 1.0, OC_X, alpha, OC_PHAT);
        rnorm = oc_norm(c, m, OC_S);
        if (hist && nh < hist_cap) hist[nh] = -rnorm;
        nh++;
        if (rnorm < rnorm0 * tol_rel || rnorm < tol_abs) { conv = 1; break; }
        if (!precond) copy_valid(c, m, OC_SHAT, OC_S);
        else oc_precond(c, m, OC_SHAT, OC_S);
        oc_apply(c, m, OC_T, OC_SHAT);
        const double ts = oc_dot(c, m, OC_T, OC_S);
        const double tt = oc_dot(c, m, OC_T, OC_T);
        if (tt == 0) { stat = 1; break; }
        omega = ts / tt;
        saxby(c, m, OC_X, 1.0, OC_X, omega, OC_SHAT);
        saxby(c, m, OC_R, 1.0, OC_S, -omega, OC_T);
        rnorm = oc_norm(c, m, OC_R);
        if (hist && nh < hist_cap) hist[nh] = rnorm;
        nh++;
        if (rnorm < rnorm0 * tol_rel + tol_abs) { conv = 1; break; }
        if (omega == 0) { stat = 1; break; }
        rho1 = rho;
    }
    if (conv) {
        correct(c, m, OC_X0, OC_X);
        copy_valid(c, m, solc, OC_X0);
    }
    if (converged) *converged = conv;
    if (nhist) *nhist = nh;
    return stat;
}

int oc_cg_level(oc_ctx* c, int m, int solc, int rhsc, double tol_rel, double tol_abs, int max_iter, int precond,
                int* iters, int* converged, double* hist, int hist_cap, int* nhist) {
    return cg_level(c, m, solc, rhsc, tol_rel, tol_abs, max_iter, precond, iters, converged, hist, hist_cap, nhist);
}

int oc_cg_solve(oc_ctx* c, int* iters, int* converged, double* hist, int hist_cap, int* nhist) {
    return cg_level(c, finest(c), OC_SOL, OC_RHS, c->tol_rel, c->tol_abs, c->max_iter, c->precond, iters, converged,
                    hist, hist_cap, nhist);
}

void oc_bottom_stats(oc_ctx* c, int reset, int* calls, long long* iters) {
    if (calls) *calls = c->bottom_calls;
    if (iters) *iters = c->bottom_iters;
    if (reset) { c->bottom_calls = 0; c->bottom_iters = 0; }
}
