// capi_tree.cpp -- extern "C" entry points of the host tree (for ctypes and other FFIs).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "amr_tree.h"
#include "fill_patch.h"
#include "sayaka_amd.h"
#include "sayaka_host.h"
#include "tree_data.h"

using namespace sayaka;

namespace {
// the registered states of each tree handle, owned here (the reference's drivers own theirs)
struct OwnedState {
    TreeData data;
    BCPatch bc;
    std::unique_ptr<FillPatch> fill;
};
std::map<const AmrTree*, std::vector<std::unique_ptr<OwnedState>>> g_states;
thread_local std::string g_terr;
}  // namespace

extern "C" {

sy_tree* sy_tree_create(int ndim, int nb, int ngrow, int minlevel, int maxlevel, int max_blocks,
                        const int* ndiv, const double* lo, const double* hi, const int* bc) {
    if ((ndim != 2 && ndim != 3) || nb < 2 || nb % 2 || minlevel < 1 || maxlevel < minlevel ||
        maxlevel >= MAX_LEVEL || max_blocks < 1)
        return nullptr;
    AmrTree* t = new (std::nothrow) AmrTree;
    if (!t) return nullptr;
    t->ndim = ndim;
    t->nb = nb;
    t->blockNumGrow = ngrow;
    t->minAmrLevel = minlevel;
    t->maxAmrLevel = maxlevel;
    t->maxBlockNum = max_blocks;
    t->init();
    RealBox pb;
    pb.vlo = Vector3d::VecMake(lo[0], lo[1], lo[2]);
    pb.vhi = Vector3d::VecMake(hi[0], hi[1], hi[2]);
    bool per[MAX_FACE] = {false, false, false, false, false, false};
    int bct[MAX_FACE];
    for (int f = 0; f < MAX_FACE; f++) {
        per[f] = bc[f] == SY_TREE_PERIODIC;
        bct[f] = per[f] ? -100 : bc[f];
    }
    const int nd[3] = {ndiv[0], ndiv[1], ndim == 3 ? ndiv[2] : 1};
    t->initUniformRootLevel(nd, pb, bct, per);
    t->initRefineToMinLevel();
    return reinterpret_cast<sy_tree*>(t);
}

void sy_tree_destroy(sy_tree* h) {
    AmrTree* t = reinterpret_cast<AmrTree*>(h);
    if (t) t->treeStateData.clear();
    g_states.erase(t);
    delete t;
}

const char* sy_tree_last_error(void) { return g_terr.c_str(); }

int sy_tree_state_register(sy_tree* h, int dir, int ncomp, int ngrow, const int* prolong, int nphys,
                           const int* physbc, const int* mathbc, const double* bcval) {
    AmrTree* t = reinterpret_cast<AmrTree*>(h);
    try {
        if (!t || !prolong || ncomp < 1 || ngrow < 0 || (nphys > 0 && (!physbc || !mathbc || !bcval)))
            throw std::invalid_argument("bad argument");
        auto st = std::make_unique<OwnedState>();
        st->data = dir < 0 ? TreeData::MakeCellData(*t, ncomp, ngrow) : TreeData::MakeFaceData(*t, dir, ncomp, ngrow);
        st->bc.defineByData(st->data);
        for (int c = 0; c < ncomp; c++)
            for (int i = -1; i <= 1; i++)
                for (int j = -1; j <= 1; j++)
                    for (int k = -1; k <= 1; k++)
                        for (int q = 0; q < nphys; q++) {
                            st->bc.boundaryRegister(c).setBCMap(i, j, k, physbc[q], mathbc[c * nphys + q]);
                            st->bc.boundaryRegister(c).setBCVal(i, j, k, physbc[q], bcval[c * nphys + q]);
                        }
        st->fill = std::make_unique<FillPatch>(*t, st->data);
        st->fill->setBC(st->bc);
        for (int c = 0; c < ncomp; c++) st->fill->setProlongation(c, prolong[c]);
        TreeStateData tsd;
        tsd.new_data = &st->data;
        tsd.bc_patch = &st->bc;
        tsd.fill_new_patch = st->fill.get();
        t->treeStateData.push_back(tsd);
        g_states[t].push_back(std::move(st));
        return (int)t->treeStateData.size() - 1;
    } catch (const std::exception& e) {
        g_terr = e.what();
        return -1;
    }
}

static TreeData* state_data(const AmrTree* t, int idx) {
    if (!t || idx < 0 || idx >= (int)t->treeStateData.size()) return nullptr;
    return t->treeStateData[idx].new_data;
}

int sy_tree_state_set_grown(sy_tree* h, int idx, const double* data) {
    TreeData* d = state_data(reinterpret_cast<AmrTree*>(h), idx);
    if (!d || !data) {
        g_terr = "bad state index";
        return -1;
    }
    const int n = reinterpret_cast<AmrTree*>(h)->numBlocks;
    const size_t per = (size_t)n * d->grownSize();
    for (int c = 0; c < d->numComp(); c++) d->unpackGrown(c, n, data + c * per);
    try {
        if (d->resident()) d->toDevice();
    } catch (const std::exception& e) {
        g_terr = e.what();
        return -1;
    }
    return 0;
}

int sy_tree_state_get_grown(const sy_tree* h, int idx, double* data) {
    TreeData* d = state_data(reinterpret_cast<const AmrTree*>(h), idx);
    if (!d || !data) {
        g_terr = "bad state index";
        return -1;
    }
    try {
        if (d->resident()) d->toHost();
    } catch (const std::exception& e) {
        g_terr = e.what();
        return -1;
    }
    const int n = reinterpret_cast<const AmrTree*>(h)->numBlocks;
    const size_t per = (size_t)n * d->grownSize();
    for (int c = 0; c < d->numComp(); c++) d->packGrown(c, n, data + c * per);
    return 0;
}

int sy_tree_state_set_resident(sy_tree* h, int idx, int on, int device) {
    TreeData* d = state_data(reinterpret_cast<AmrTree*>(h), idx);
    if (!d) {
        g_terr = "bad state index";
        return -1;
    }
    try {
        d->setResident(on != 0, device);
        return 0;
    } catch (const std::exception& e) {
        g_terr = e.what();
        return -1;
    }
}

int sy_tree_state_fill(sy_tree* h, int idx, int scomp, int ncomp, int nlayer) {
    AmrTree* t = reinterpret_cast<AmrTree*>(h);
    if (!state_data(t, idx) || !t->treeStateData[idx].fill_new_patch) {
        g_terr = "bad state index";
        return -1;
    }
    try {
        t->treeStateData[idx].fill_new_patch->fillBoundary(scomp, ncomp, nlayer);
        return 0;
    } catch (const std::exception& e) {
        g_terr = e.what();
        return -1;
    }
}

int sy_tree_transfer_stats(const sy_tree* h, long long* out4) {
    const AmrTree* t = reinterpret_cast<const AmrTree*>(h);
    if (!t || !out4) {
        g_terr = "bad argument";
        return -1;
    }
    try {
        return sayaka_dev_transfer_stats(t->deviceContext(0), out4) < 0 ? -1 : 0;
    } catch (const std::exception& e) {
        g_terr = e.what();
        return -1;
    }
}

int sy_tree_init_new_data(sy_tree* h) {
    try {
        reinterpret_cast<AmrTree*>(h)->initNewBlockDataAfterRegrid();
        return 0;
    } catch (const std::exception& e) {
        g_terr = e.what();
        return -1;
    }
}

int sy_tree_numblocks(const sy_tree* h) { return reinterpret_cast<const AmrTree*>(h)->numBlocks; }
int sy_tree_finest(const sy_tree* h) { return reinterpret_cast<const AmrTree*>(h)->currentFinestLevel(); }

int sy_tree_set_flags(sy_tree* h, const int* refine, const int* coarsen, int n) {
    AmrTree* t = reinterpret_cast<AmrTree*>(h);
    if (n < 0 || n > t->maxTreeBlockNum) return -1;
    for (int i = 0; i < n; i++) {
        if (refine) t->refineFlag[i] = refine[i];
        if (coarsen) t->coarsenFlag[i] = coarsen[i];
    }
    return 0;
}

int sy_tree_get_flags(const sy_tree* h, int* refine, int* coarsen, int n) {
    const AmrTree* t = reinterpret_cast<const AmrTree*>(h);
    if (n < 0 || n > t->maxTreeBlockNum) return -1;
    for (int i = 0; i < n; i++) {
        if (refine) refine[i] = t->refineFlag[i];
        if (coarsen) coarsen[i] = t->coarsenFlag[i];
    }
    return 0;
}

int sy_tree_regrid(sy_tree* h) {
    AmrTree* t = reinterpret_cast<AmrTree*>(h);
    return t->regrid() == 0 ? t->numBlocks : -1;
}

void sy_tree_regrid_origin(const sy_tree* h, int* origin) {
    const AmrTree* t = reinterpret_cast<const AmrTree*>(h);
    for (int i = 0; i < t->numBlocks; i++) origin[i] = i < (int)t->regridOrigin.size() ? t->regridOrigin[i] : i;
}

void sy_tree_export(const sy_tree* h, int* level, int* parent, int* which_child, int* child8, int* nbr6,
                    int* node_type, int* node_type2, double* lo3, double* hi3, double* center3,
                    double* length3, int* surr27, int* surrtype27, int* surrbcref27) {
    const AmrTree* t = reinterpret_cast<const AmrTree*>(h);
    for (int b = 0; b < t->numBlocks; b++) {
        const AmrTreeNode& n = t->blocks[b];
        level[b] = n.refineLevel;
        parent[b] = n.parent;
        which_child[b] = n.whichChild;
        node_type[b] = n.nodeType;
        std::memcpy(child8 + 8 * b, n.child, 8 * sizeof(int));
        std::memcpy(nbr6 + 6 * b, n.neighbor, 6 * sizeof(int));
        std::memcpy(node_type2 + 6 * b, n.nodeType2, 6 * sizeof(int));
        for (int d = 0; d < 3; d++) {
            lo3[3 * b + d] = n.boundBox.vlo(d);
            hi3[3 * b + d] = n.boundBox.vhi(d);
            center3[3 * b + d] = n.blockCenter(d);
            length3[3 * b + d] = n.blockLength(d);
        }
        const SurroundingBlocks& s = t->surrBlocks[b];
        std::memcpy(surr27 + 27 * b, &s.blocks[0][0][0], 27 * sizeof(int));
        std::memcpy(surrtype27 + 27 * b, &s.types[0][0][0], 27 * sizeof(int));
        std::memcpy(surrbcref27 + 27 * b, &s.bcrefs[0][0][0], 27 * sizeof(int));
    }
}

// Does some cell centre of block n lie within `half` of the shell |x - centre| = r0?  Exactly
// the per-cell test (|sqrt(r2) - r0| < half over the centres, i fastest), decided without the
// scan where it cannot go either way: the centres' distances run from the nearest centre's to
// the farthest's in steps of at most one cell width (walk one axis at a time), so when that range
// misses the band by a margin no centre can hit, and when it overlaps the band by more than a
// cell width some centre must.  The 1e-12 margin covers the rounding of the distances.
static bool shell_hit(const AmrTree& t, const AmrTreeNode& n, const double* centre, double r0, double half) {
    const int nb = t.nb, nz = t.ndim == 3 ? nb : 1;
    double dx[3];
    for (int d = 0; d < 3; d++) dx[d] = n.blockLength(d) / nb;
    auto dist = [&](int i, int j, int k) {
        const double x = n.boundBox.vlo(0) + (i + 0.5) * dx[0];
        const double y = n.boundBox.vlo(1) + (j + 0.5) * dx[1];
        double r2 = (x - centre[0]) * (x - centre[0]) + (y - centre[1]) * (y - centre[1]);
        if (t.ndim == 3) {
            const double z = n.boundBox.vlo(2) + (k + 0.5) * dx[2];
            r2 += (z - centre[2]) * (z - centre[2]);
        }
        return std::sqrt(r2);
    };
    int inear[3] = {0, 0, 0}, ifar[3] = {0, 0, 0};
    double step = 0.0;
    for (int d = 0; d < t.ndim; d++) {
        const int cnt = d == 2 ? nz : nb;
        const double u = (centre[d] - n.boundBox.vlo(d)) / dx[d] - 0.5;
        inear[d] = std::min(cnt - 1, std::max(0, (int)std::lround(u)));
        ifar[d] = u < 0.5 * (cnt - 1) ? cnt - 1 : 0;
        step = std::max(step, dx[d]);
    }
    const double eps = 1e-12 * std::max(1.0, r0 + half);
    const double dmin = dist(inear[0], inear[1], inear[2]), dmax = dist(ifar[0], ifar[1], ifar[2]);
    if (dmin - r0 >= half + eps || r0 - dmax >= half + eps) return false;
    const double lo = std::max(dmin, r0 - half + eps), hi = std::min(dmax, r0 + half - eps);
    if (hi - lo > step * (1.0 + 1e-12)) return true;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < nb; j++)
            for (int i = 0; i < nb; i++)
                if (std::fabs(dist(i, j, k) - r0) < half) return true;
    return false;
}

// Refinement around a spherical shell, level by level from minAmrLevel: a leaf (or parent of
// leaves) of level l < maxAmrLevel is flagged when one of its cell centres
// x = lo + (i + 1/2) * dx lies within factor * blockLength of |x - c| = r0; one regrid per level
// plus a final one (the same procedure as workloads.build_tree with shell_criterion).
int sy_tree_build_shell(sy_tree* h, const double* centre, double r0, double factor) {
    AmrTree* t = reinterpret_cast<AmrTree*>(h);
    for (int level = t->minAmrLevel; level <= t->maxAmrLevel; level++) {
        std::fill(t->refineFlag.begin(), t->refineFlag.end(), 0);
        if (level < t->maxAmrLevel)
            for (int b = 0; b < t->numBlocks; b++) {
                const AmrTreeNode& n = t->blocks[b];
                if (n.refineLevel != level || !(n.nodeType == NodeIsLeaf || n.nodeType == NodeHasLeafChild)) continue;
                t->refineFlag[b] = shell_hit(*t, n, centre, r0, factor * n.blockLength(0)) ? 1 : 0;
            }
        if (t->regrid() != 0) return -1;
    }
    return t->numBlocks;
}

// Refine / coarsen flags for the shell |x - centre| = r0 on the current tree: a leaf is hit
// when one of its cell centres lies within factor * (its block edge) of the shell (the rule
// of sy_tree_build_shell); hit leaves below maxAmrLevel are flagged for refinement, leaves
// that are not hit and above minAmrLevel for coarsening (regrid keeps only complete sibling
// sets and the 2:1 balance).  Blocks whose box cannot reach the band are rejected first;
// blocks run in parallel.  Returns the number of flagged blocks.
int sy_tree_flag_shell(sy_tree* h, const double* centre, double r0, double factor) {
    AmrTree* t = reinterpret_cast<AmrTree*>(h);
    std::fill(t->refineFlag.begin(), t->refineFlag.end(), 0);
    std::fill(t->coarsenFlag.begin(), t->coarsenFlag.end(), 0);
    int nflag = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : nflag)
    for (int b = 0; b < t->numBlocks; b++) {
        const AmrTreeNode& n = t->blocks[b];
        if (n.nodeType != NodeIsLeaf) continue;
        const double half = factor * n.blockLength(0);
        double dmin2 = 0.0, dmax2 = 0.0;
        for (int d = 0; d < t->ndim; d++) {
            const double lo = n.boundBox.vlo(d) - centre[d], hi = n.boundBox.vhi(d) - centre[d];
            const double near = lo > 0 ? lo : (hi < 0 ? -hi : 0.0);
            const double far = std::max(std::fabs(lo), std::fabs(hi));
            dmin2 += near * near;
            dmax2 += far * far;
        }
        const bool hit = std::sqrt(dmin2) <= r0 + half && std::sqrt(dmax2) >= r0 - half &&
                         shell_hit(*t, n, centre, r0, half);
        if (hit && n.refineLevel < t->maxAmrLevel) {
            t->refineFlag[b] = 1;
            nflag++;
        } else if (!hit && n.refineLevel > t->minAmrLevel) {
            t->coarsenFlag[b] = 1;
            nflag++;
        }
    }
    return nflag;
}

int sy_tree_level_count(const sy_tree* h, int l) {
    const AmrTree* t = reinterpret_cast<const AmrTree*>(h);
    return (l >= 1 && l <= t->maxAmrLevel) ? t->levels[l].numLevelBlock : 0;
}
void sy_tree_level_blocks(const sy_tree* h, int l, int* out) {
    const AmrTree* t = reinterpret_cast<const AmrTree*>(h);
    if (l < 1 || l > t->maxAmrLevel) return;
    std::memcpy(out, t->levels[l].levelBlock.data(), t->levels[l].levelBlock.size() * sizeof(int));
}

}  // extern "C"
