// amr_tree.cpp -- regrid of the host-side tree.
//
// The numbering of nodes is part of the contract with every caller that keeps per-block
// data (TreeData, the device block arrays, the solver's index tables), so each step below
// produces exactly the ids, neighbour links and node types of the reference regrid
// (SayakaRegrid.cpp:26-976, SayakaTree.cpp:34-233).  The comments name the rule each
// loop implements; the loops are written against this file's run-time dimension.
#include "amr_tree.h"

#include "dev_tree.h"
#include "fill_patch.h"
#include "sayaka_amd.h"
#include "tree_data.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

namespace sayaka {
inline namespace amd {

void TreeDevice::release() {
    if (ctx) sayaka_dev_destroy(ctx);
    ctx = nullptr;
    device = -1;
    topo_version = surr_version = ~0ull;
}

sayaka_dev_ctx* AmrTree::deviceContext(int device) const {
    if (device_.ctx) {
        if (device != device_.device)
            throw std::invalid_argument("AmrTree: the tree's device context is on GPU " +
                                        std::to_string(device_.device) + ", not " + std::to_string(device));
        return device_.ctx;
    }
    sayaka_dev_ctx* c = nullptr;
    if (sayaka_dev_create(device, ndim, nb, maxBlockNum, &c) < 0)
        throw std::runtime_error(std::string("AmrTree: device context: ") + sayaka_dev_last_error());
    device_.ctx = c;
    device_.device = device;
    return c;
}

sayaka_dev_ctx* AmrTree::deviceTables(int device, bool surroundings) const {
    sayaka_dev_ctx* c = deviceContext(device);
    if (device_.topo_version != topologyVersion) {
        upload_tree_topology(c, *this);
        device_.topo_version = topologyVersion;
        device_.surr_version = ~0ull;   // the surrounding tables follow the block count
    }
    if (surroundings && device_.surr_version != topologyVersion) {
        upload_tree_surroundings(c, *this);
        device_.surr_version = topologyVersion;
    }
    return c;
}

void AmrTreeNode::setUnused() {
    refineLevel = -1;
    std::fill(neighbor, neighbor + MAX_FACE, -1);
    std::fill(child, child + MAX_CHILD, -1);
    parent = -1;
    whichChild = -1;
    nodeType = -1;
    std::fill(nodeType2, nodeType2 + MAX_FACE, -1);
    blockFlags = -1;
    blockCenter.setZero();
    blockLength.setZero();
    boundBox.vlo.setZero();
    boundBox.vhi.setZero();
}

void SurroundingBlocks::setUnused() {
    for (int q = 0; q < 27; q++) {
        (&blocks[0][0][0])[q] = -1;
        (&types[0][0][0])[q] = -1;
        (&bcrefs[0][0][0])[q] = -1;
    }
}

// Index of a surrounding position, z term only in 3D (SayakaSurroundingIndex.h:69-81).
int AmrTree::surrIndex(int i, int j, int k) const {
    int m = i + 1 + (j + 1) * 3;
    if (ndim == 3) m += (k + 1) * 9;
    return m;
}

// Pool allocation (SayakaTree.cpp:34-91).  A caller that set the reference's index boxes
// gets ndim and nb from them; otherwise the boxes are made from ndim, nb and blockNumGrow.
void AmrTree::init() {
    if (blockValidIndex.size(0) > 1) {
        nb = blockValidIndex.size(0);
        ndim = blockValidIndex.size(2) > 1 ? 3 : 2;
    } else {
        blockValidIndex = IndexBox(Vector3i{0, 0, 0}, Vector3i{nb - 1, nb - 1, ndim == 3 ? nb - 1 : 0});
    }
    if (blockGrownIndex.size(0) <= blockValidIndex.size(0)) {
        blockGrownIndex = blockValidIndex;
        blockGrownIndex.extend(blockNumGrow, blockNumGrow, ndim == 3 ? blockNumGrow : 0);
    }
    maxTreeBlockNum = maxBlockNum;
    numBlocks = 0;
    gridChanged = 0;
    topologyVersion++;
    device_ = TreeDevice{};   // the pool size may have changed: a new context on first use
    blocks.assign(maxTreeBlockNum, AmrTreeNode{});
    for (auto& b : blocks) b.setUnused();
    refineFlag.assign(maxTreeBlockNum, 0);
    coarsenFlag.assign(maxTreeBlockNum, 0);
    newChildFlag.assign(maxTreeBlockNum, 0);
    stayFlag.assign(maxTreeBlockNum, 1);
    surrBlocks.assign(maxTreeBlockNum, SurroundingBlocks{});
    for (auto& s : surrBlocks) s.setUnused();
    levels.assign(maxAmrLevel + 1, AmrTreeLevel{});
    for (int l = 0; l <= maxAmrLevel; l++) levels[l].level = l;
}

// Root level: n0x*n0y*n0z leaves at level 1, i fastest; walls get the caller's boundary
// code, periodic faces wrap (SayakaTree.cpp:93-201).
void AmrTree::initUniformRootLevel(const int ndiv[3], const RealBox& pb, const int bc_type[MAX_FACE],
                                   const bool is_periodic[MAX_FACE]) {
    const int n0x = ndiv[0], n0y = ndiv[1], n0z = ndiv[2];
    auto ind = [=](int i, int j, int k) { return i + j * n0x + k * n0x * n0y; };
    numBlocks = n0x * n0y * n0z;
    for (int k = 0; k < n0z; k++)
        for (int j = 0; j < n0y; j++)
            for (int i = 0; i < n0x; i++) {
                AmrTreeNode& r = blocks[ind(i, j, k)];
                Vector3d lo = pb.vlo, hi = pb.vlo;
                lo.x += pb.length(0) / n0x * i;
                hi.x += pb.length(0) / n0x * (i + 1);
                lo.y += pb.length(1) / n0y * j;
                hi.y += pb.length(1) / n0y * (j + 1);
                if (ndim == 3) {
                    lo.z += pb.length(2) / n0z * k;
                    hi.z += pb.length(2) / n0z * (k + 1);
                }
                r.boundBox.vlo = lo;
                r.boundBox.vhi = hi;
                for (int d = 0; d < 3; d++) {
                    r.blockCenter(d) = 0.5 * (lo(d) + hi(d));
                    r.blockLength(d) = hi(d) - lo(d);
                }
                r.nodeType = NodeIsLeaf;
                r.refineLevel = 1;
                const int pos[3] = {i, j, k};
                const int cnt[3] = {n0x, n0y, n0z};
                for (int f = 0; f < numFace(); f++) {
                    const int d = face_dir(f), s = face_side(f);
                    int q[3] = {i, j, k};
                    int nbr;
                    const bool at_wall = s == 0 ? pos[d] == 0 : pos[d] == cnt[d] - 1;
                    if (at_wall) {
                        q[d] = s == 0 ? cnt[d] - 1 : 0;
                        nbr = is_periodic[f] ? ind(q[0], q[1], q[2]) : bc_type[f];
                    } else {
                        q[d] += s == 0 ? -1 : 1;
                        nbr = ind(q[0], q[1], q[2]);
                    }
                    r.neighbor[f] = nbr;
                }
            }
    topologyVersion++;
}

// Regrid once per level up to minAmrLevel; regrid() itself flags every leaf below
// minAmrLevel (SayakaTree.cpp:204-233).
void AmrTree::initRefineToMinLevel() {
    for (int l = 1; l <= minAmrLevel; l++) regrid();
}

Vector3d AmrTree::getBlockCellSize(int b) const {
    Vector3d s = blocks[b].blockLength;
    for (int d = 0; d < ndim; d++) s(d) /= nb;
    return s;
}
double AmrTree::getBlockCellVolume(int b) const {
    Vector3d dx = getBlockCellSize(b);
    double v = 1.0;
    for (int d = 0; d < ndim; d++) v *= dx(d);
    return v;
}
Vector3d AmrTree::getBlockCellArea(int b) const {
    Vector3d dx = getBlockCellSize(b);
    const double v = getBlockCellVolume(b);
    Vector3d ds = Vector3d::VecMake(1.0, 1.0, 1.0);
    for (int d = 0; d < ndim; d++) ds(d) = v / dx(d);
    return ds;
}
int AmrTree::currentFinestLevel() const {
    int f = 0;
    for (int i = 0; i < numBlocks; i++) f = std::max(f, blocks[i].refineLevel);
    return f;
}

// ------------------------------------------------------------------------------ regrid
// Flag clamping, coarsen checks, refine then pack (SayakaRegrid.cpp:26-125).
int AmrTree::regrid() {
    std::vector<int> refine_save(maxTreeBlockNum, 0);
    for (int i = 0; i < numBlocks; i++) {
        const int l = blocks[i].refineLevel;
        if (l >= maxAmrLevel) refineFlag[i] = 0;
        if (l <= minAmrLevel) coarsenFlag[i] = 0;
        if (l < minAmrLevel && blocks[i].isLeaf()) refineFlag[i] = 1;
        refine_save[i] = refineFlag[i];
    }
    std::fill(newChildFlag.begin(), newChildFlag.end(), 0);
    regridOrigin.assign(maxTreeBlockNum, -1);
    for (int i = 0; i < numBlocks; i++) regridOrigin[i] = i;
    checkCoarsen();
    // a leaf is not coarsened when its parent is being refined
    for (int i = 0; i < numBlocks; i++)
        if (blocks[i].isLeaf() && coarsenFlag[i] && refine_save[blocks[i].parent]) coarsenFlag[i] = 0;
    bool need = false;
    for (int i = 0; i < numBlocks && !need; i++)
        if (blocks[i].isLeaf() && (refineFlag[i] || coarsenFlag[i])) need = true;
    if (need) {
        int nrefine = 0;
        for (int i = 0; i < numBlocks; i++)
            if (blocks[i].isLeaf() && refineFlag[i]) nrefine++;
        int nsave = numBlocks;
        if (nrefine > 0 && refineBlocks() != 0) return -1;
        nsave = coarsenBlocks(nsave);
        topologyVersion++;
    }
    buildSurroundingBlocks();
    cacheTreeLevels();
    gridChanged = 1;
    for (TreeStateData& st : treeStateData)
        if (st.new_data) st.new_data->remapBlocks(regridOrigin, numBlocks);
    return 0;
}

void AmrTree::initNewBlockDataAfterRegrid() {
    bool any = false;
    for (int i = 0; i < numBlocks && !any; i++) any = regridOrigin.size() > (size_t)i && regridOrigin[i] < 0;
    if (!any) return;   // SayakaRegrid.cpp:1004-1005
    std::vector<int> isnew(numBlocks);
    for (int i = 0; i < numBlocks; i++) isnew[i] = regridOrigin[i] < 0;
    for (TreeStateData& st : treeStateData)
        if (st.new_data && st.fill_new_patch) st.fill_new_patch->initNewBlocks(isnew);
}

// A leaf may be coarsened only when no face neighbour has children, no neighbour of such a
// block (edge reach, 2D/3D) and, in 3D, no neighbour of those (corner reach) does; roots and
// incomplete sibling sets stay (SayakaRegrid.cpp:127-238).
void AmrTree::checkCoarsen() {
    const int M = maxTreeBlockNum;
    std::vector<int> nonleaf(M, 0), t1(M, 0), t2(M, 0), t3(M, 0);
    for (int i = 0; i < numBlocks; i++)
        if (coarsenFlag[i] && !blocks[i].isLeaf()) coarsenFlag[i] = 0;
    for (int i = 0; i < numBlocks; i++) nonleaf[i] = blocks[i].hasChildren() ? 1 : 0;
    auto spread = [&](const std::vector<int>& src, std::vector<int>& dst) {
        const int n = numBlocks;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; i++)
            for (int f = 0; f < numFace(); f++) {
                const int n = blocks[i].neighbor[f];
                if (n >= 0 && src[n]) { dst[i] = 1; break; }
            }
    };
    spread(nonleaf, t1);
    if (ndim >= 2) {
        spread(t1, t2);
        for (int i = 0; i < numBlocks; i++) t2[i] = (t1[i] || t2[i]) ? 1 : 0;
    }
    if (ndim == 3) spread(t2, t3);
    for (int i = 0; i < numBlocks; i++)
        if (blocks[i].isLeaf() && (t1[i] || t2[i] || t3[i])) coarsenFlag[i] = 0;
    for (int i = 0; i < numBlocks; i++)
        if (blocks[i].parent < 0 && coarsenFlag[i]) coarsenFlag[i] = 0;
    for (int i = 0; i < numBlocks; i++) {
        if (!coarsenFlag[i]) continue;
        const AmrTreeNode& p = blocks[blocks[i].parent];
        for (int c = 0; c < numChild(); c++)
            if (!coarsenFlag[p.child[c]]) { coarsenFlag[i] = 0; break; }
    }
}

// Refine flagged leaves, link the new children, then repeat with the blocks that the 2:1
// balance rules flag until none are left (SayakaRegrid.cpp:240-510).
int AmrTree::refineBlocks() {
    const int M = maxTreeBlockNum;
    std::vector<int> t1(M, 0), t2(M, 0), t3(M, 0);
    std::fill(newChildFlag.begin(), newChildFlag.end(), 0);
    const int NF = numFace(), NC = numChild();
    // The per-block passes below write only block i's own entries from values that earlier
    // passes wrote: they run in parallel with the sequential result.  Child creation and
    // linking (ids handed out in order, links written across blocks) stay sequential.
    for (int cycle = 0; cycle < 99999; cycle++) {
        const int nb0 = numBlocks;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nb0; i++)
            if (refineFlag[i] && !blocks[i].isLeaf()) refineFlag[i] = 0;

        // children take the next free ids, parent by parent, child index order
        int nblock = numBlocks;
        for (int i = 0; i < numBlocks; i++) {
            if (!refineFlag[i]) continue;
            for (int c = 0; c < NC; c++) {
                if (nblock >= M) {
                    std::fprintf(stderr, "AmrTree::regrid: block pool overflow (%d)\n", M);
                    return -1;
                }
                const int j = nblock++;
                AmrTreeNode& p = blocks[i];
                AmrTreeNode& ch = blocks[j];
                p.child[c] = j;
                ch.refineLevel = p.refineLevel + 1;
                ch.parent = i;
                ch.whichChild = c;
                ch.boundBox = p.boundBox;
                for (int d = 0; d < ndim; d++) {
                    if (child_side(c, d) == 0) ch.boundBox.vhi(d) = p.boundBox.center(d);
                    else ch.boundBox.vlo(d) = p.boundBox.center(d);
                }
                for (int d = 0; d < 3; d++) {
                    ch.blockCenter(d) = 0.5 * (ch.boundBox.vlo(d) + ch.boundBox.vhi(d));
                    ch.blockLength(d) = ch.boundBox.vhi(d) - ch.boundBox.vlo(d);
                }
                ch.blockFlags = p.blockFlags;
                newChildFlag[j] = 1;
            }
        }
        // neighbours of the new children: siblings directly, otherwise the matching child
        // of the parent's neighbour, which is also linked back
        for (int i = 0; i < numBlocks; i++) {
            if (!refineFlag[i]) continue;
            for (int c = 0; c < NC; c++) {
                const int ic = blocks[i].child[c];
                for (int f = 0; f < NF; f++) {
                    bool sib;
                    const int cc = child_neighbor(c, f, sib);
                    if (sib) {
                        blocks[ic].neighbor[f] = blocks[i].child[cc];
                    } else {
                        const int pn = blocks[i].neighbor[f];
                        if (pn >= 0) {
                            const int jc = blocks[pn].child[cc];
                            blocks[ic].neighbor[f] = jc;
                            if (jc >= 0) blocks[jc].neighbor[face_opposite(f)] = ic;
                        }
                    }
                }
            }
        }
        numBlocks = nblock;

        // face marks: 2 on every face of a parent, 3 on face (d, 1-side_d(c)) of a parent
        // whose child c has children of its own (SayakaRegrid.cpp:350-382)
        const int nb1 = numBlocks;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nb1; i++) {
            int* t2f = blocks[i].nodeType2;
            for (int f = 0; f < NF; f++) t2f[f] = 1;
            for (int c = 0; c < NC; c++) {
                const int j = blocks[i].child[c];
                if (j < 0) continue;
                for (int f = 0; f < NF; f++)
                    if (t2f[f] != 3) t2f[f] = 2;
                for (int cc = 0; cc < NC; cc++)
                    if (blocks[j].child[cc] >= 0)
                        for (int d = 0; d < ndim; d++) t2f[2 * d + 1 - child_side(c, d)] = 3;
            }
        }
        setNodeType();

        std::fill(t1.begin(), t1.end(), 0);
        std::fill(t2.begin(), t2.end(), 0);
        std::fill(t3.begin(), t3.end(), 0);
        std::fill(refineFlag.begin(), refineFlag.end(), 0);
        // rule 1: a face neighbour whose mark on that same face index is 3
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nb1; i++)
            for (int f = 0; f < NF; f++) {
                const int n = blocks[i].neighbor[f];
                if (n >= 0 && blocks[n].nodeType2[f] == 3) { t1[i] = 1; break; }
            }
        // rule 2 (edges): rule-1 neighbours in at least two directions; rule 3 (3D corners):
        // rule-1/2 neighbours in all three directions
        auto dirs_hit = [&](int i, const std::vector<int>& src, int cnt[3]) {
            cnt[0] = cnt[1] = cnt[2] = 0;
            for (int f = 0; f < NF; f++) {
                const int n = blocks[i].neighbor[f];
                if (n >= 0 && src[n]) cnt[face_dir(f)]++;
            }
        };
        if (ndim >= 2) {
#pragma omp parallel for schedule(static)
            for (int i = 0; i < nb1; i++) {
                int c[3];
                dirs_hit(i, t1, c);
                if ((c[0] >= 1 && c[1] >= 1) || (c[0] >= 1 && c[2] >= 1) || (c[1] >= 1 && c[2] >= 1)) t2[i] = 1;
            }
#pragma omp parallel for schedule(static)
            for (int i = 0; i < nb1; i++)
                if (t2[i] || t1[i]) t2[i] = 1;
        }
        if (ndim == 3) {
#pragma omp parallel for schedule(static)
            for (int i = 0; i < nb1; i++) {
                int c[3];
                dirs_hit(i, t2, c);
                if (c[0] >= 1 && c[1] >= 1 && c[2] >= 1) t3[i] = 1;
            }
        }
        int repeat = 0;
#pragma omp parallel for schedule(static) reduction(| : repeat)
        for (int i = 0; i < nb1; i++) {
            refineFlag[i] = 0;
            if (blocks[i].isLeaf() && (t1[i] || t2[i] || t3[i])) {
                repeat = 1;
                refineFlag[i] = 1;
            }
        }
        if (!repeat) break;
    }
    // new children on the domain boundary inherit the parent's wall code
    for (int i = 0; i < numBlocks; i++) {
        if (!newChildFlag[i]) continue;
        const int p = blocks[i].parent;
        for (int f = 0; f < NF; f++)
            if (blocks[i].neighbor[f] < 0 && blocks[p].neighbor[f] <= PHYS_BNDRY)
                blocks[i].neighbor[f] = blocks[p].neighbor[f];
    }
    std::fill(refineFlag.begin(), refineFlag.end(), 0);
    return 0;
}

// Drop complete sibling sets flagged for coarsening, pack the survivors to the front of
// the pool keeping their order, and renumber every link (SayakaRegrid.cpp:515-632).
int AmrTree::coarsenBlocks(int oldBlockNum) {
    int nold = oldBlockNum;
    checkCoarsen();
    std::vector<int> newloc(maxTreeBlockNum, -1);
    for (int i = 0; i < numBlocks; i++)
        if (coarsenFlag[i]) blocks[i].blockFlags = -1;
    int nblock = numBlocks;
    int ipack = 0;
    for (int i = 0; i < numBlocks; i++) {
        if (!coarsenFlag[i]) {
            if (ipack != i) {
                blocks[ipack] = blocks[i];
                newChildFlag[ipack] = newChildFlag[i];
                regridOrigin[ipack] = regridOrigin[i];
            }
            newloc[i] = ipack++;
        } else {
            nblock--;
            nold--;
        }
    }
    for (int i = nblock; i < numBlocks; i++) {
        coarsenFlag[i] = 0;
        blocks[i].setUnused();
    }
    numBlocks = nblock;
    for (int i = 0; i < numBlocks; i++) {
        AmrTreeNode& b = blocks[i];
        if (b.parent >= 0) b.parent = newloc[b.parent];
        for (int c = 0; c < numChild(); c++)
            if (b.child[c] >= 0) b.child[c] = newloc[b.child[c]];
        for (int f = 0; f < numFace(); f++)
            if (b.neighbor[f] >= 0) b.neighbor[f] = newloc[b.neighbor[f]];
    }
    setNodeType();
    std::fill(coarsenFlag.begin(), coarsenFlag.end(), 0);
    return nold;
}

// Leaf / parent of leaves / parent of parents (SayakaRegrid.cpp:634-653).
void AmrTree::setNodeType() {
    const int n = numBlocks;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        AmrTreeNode& b = blocks[i];
        b.nodeType = NodeNoLeafChild;
        for (int c = 0; c < numChild(); c++) {
            const int j = b.child[c];
            if (j < 0) b.nodeType = NodeIsLeaf;
            else if (!blocks[j].hasChildren()) b.nodeType = NodeHasLeafChild;
        }
    }
}

// 3^d neighbourhood: faces from the node's own links, edges through a face neighbour (the
// second route wins when it reaches a same-level block), 3D corners through the three
// edge neighbours with precedence same level > coarse > wall (SayakaRegrid.cpp:655-945,
// SayakaTree.h:450-523).
void AmrTree::buildSurroundingBlocks() {
    const int ZD = ndim == 3 ? 1 : 0;
    const int center = surrIndex(0, 0, 0);
    // every block's list depends on the neighbour links only: the blocks run in parallel
#pragma omp parallel for schedule(static)
    for (int i = 0; i < numBlocks; i++) {
        SurroundingBlocks& s = surrBlocks[i];
        s.setUnused();
        s.at(0, 0, 0) = i;
        for (int f = 0; f < numFace(); f++) {
            int p[3] = {0, 0, 0};
            p[face_dir(f)] = face_side(f) ? 1 : -1;
            s.at(p[0], p[1], p[2]) = blocks[i].neighbor[f];
            s.bcrefs[p[0] + 1][p[1] + 1][p[2] + 1] = center;
        }
        for (int ks = -ZD; ks <= ZD; ks++)
            for (int js = -1; js <= 1; js++)
                for (int is = -1; is <= 1; is++) {
                    if (std::abs(is) + std::abs(js) + std::abs(ks) != 2) continue;
                    int f0, f1, a0[3], a1[3];
                    if (is == 0) {
                        f0 = js == -1 ? 2 : 3;
                        a0[0] = is; a0[1] = 0; a0[2] = ks;
                        f1 = ks == -1 ? 4 : 5;
                        a1[0] = is; a1[1] = js; a1[2] = 0;
                    } else if (js == 0) {
                        f0 = is == -1 ? 0 : 1;
                        a0[0] = 0; a0[1] = js; a0[2] = ks;
                        f1 = ks == -1 ? 4 : 5;
                        a1[0] = is; a1[1] = js; a1[2] = 0;
                    } else {
                        f0 = is == -1 ? 0 : 1;
                        a0[0] = 0; a0[1] = js; a0[2] = ks;
                        f1 = js == -1 ? 2 : 3;
                        a1[0] = is; a1[1] = 0; a1[2] = ks;
                    }
                    int& dst = s.at(is, js, ks);
                    int& ref = s.bcrefs[is + 1][js + 1][ks + 1];
                    int how = 0;  // 0 none, 1 same level, 2 coarse
                    const int b0 = s.at(a0[0], a0[1], a0[2]);
                    if (b0 >= 0) {
                        how = 1;
                        dst = blocks[b0].neighbor[f0];
                        ref = surrIndex(a0[0], a0[1], a0[2]);
                    } else if (b0 <= PHYS_BNDRY) {
                        dst = b0;
                        ref = surrIndex(a0[0], a0[1], a0[2]);
                    } else {
                        how = 2;
                        ref = surrIndex(a0[0], a0[1], a0[2]);
                    }
                    const int b1 = s.at(a1[0], a1[1], a1[2]);
                    if (b1 >= 0) {
                        dst = blocks[b1].neighbor[f1];
                        ref = surrIndex(a1[0], a1[1], a1[2]);
                    } else if (b1 <= PHYS_BNDRY) {
                        if (how == 0) {
                            dst = b1;
                            ref = surrIndex(a1[0], a1[1], a1[2]);
                        }
                    } else {
                        if (how == 0) ref = surrIndex(a1[0], a1[1], a1[2]);
                    }
                }
        if (ndim == 3) {
            for (int kk = -1; kk <= 1; kk += 2)
                for (int jj = -1; jj <= 1; jj += 2)
                    for (int ii = -1; ii <= 1; ii += 2) {
                        const int ff[3] = {ii == -1 ? 0 : 1, jj == -1 ? 2 : 3, kk == -1 ? 4 : 5};
                        int first = -1, second = -1, how = -1;  // -1 none, 0 wall, 1 same, 2 coarse
                        for (int d = 0; d < 3; d++) {
                            const int e[3] = {d == 0 ? 0 : ii, d == 1 ? 0 : jj, d == 2 ? 0 : kk};
                            const int lb = s.at(e[0], e[1], e[2]);
                            if (lb >= 0) {
                                how = 1;
                                first = blocks[lb].neighbor[ff[d]];
                                second = surrIndex(e[0], e[1], e[2]);
                            } else if (lb <= PHYS_BNDRY) {
                                if (how == -1) {
                                    how = 0;
                                    first = lb;
                                    second = surrIndex(e[0], e[1], e[2]);
                                }
                            } else if (how != 1) {
                                how = 2;
                                first = -1;
                                second = surrIndex(e[0], e[1], e[2]);
                            }
                        }
                        s.at(ii, jj, kk) = first;
                        s.bcrefs[ii + 1][jj + 1][kk + 1] = second;
                    }
        }
        for (int kk = -ZD; kk <= ZD; kk++)
            for (int jj = -1; jj <= 1; jj++)
                for (int ii = -1; ii <= 1; ii++) {
                    const int b = s.at(ii, jj, kk);
                    s.types[ii + 1][jj + 1][kk + 1] = b >= 0 ? blocks[b].nodeType : -1;
                }
    }
}

// Block ids per level in ascending order (SayakaRegrid.cpp:962-976).
void AmrTree::cacheTreeLevels() {
    for (int l = 1; l <= maxAmrLevel; l++) levels[l].clear();
    for (int i = 0; i < numBlocks; i++) levels[blocks[i].refineLevel].addBlock(i);
}

}  // namespace amd
}  // namespace sayaka
