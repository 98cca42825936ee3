// amr_tree.h -- host-side block-structured quadtree/octree (the caller-owned tree).
//
// Same call surface and same node numbering as the reference AmrTree (SayakaTree.h:195-333):
// a fixed pool of nodes, children appended at the end of the pool on refinement, survivors
// packed on coarsening, iterative 2:1 balance, a 3^d surrounding-block list per node and a
// per-level block cache.  The dimension is a run-time member (2 or 3) instead of the
// reference's compile-time SAYAKA_SPACEDIM.
#pragma once

#include <vector>

#include "index_box.h"
#include "sayaka_types.h"

struct sayaka_dev_ctx;   // a device context of the C ABI (sayaka_amd.h)

namespace sayaka {
inline namespace amd {

// The device side of a tree: one context per tree (on one GPU), shared by every MacSolver,
// FillPatch and device-resident TreeData made on it, created on first use and destroyed with
// the tree; and the topology versions its tables were last uploaded for.  A copied or
// assigned tree starts without one (its solvers and states make a new one on first use).
class TreeDevice {
   public:
    TreeDevice() = default;
    TreeDevice(const TreeDevice&) {}
    TreeDevice& operator=(const TreeDevice&) {
        release();
        return *this;
    }
    ~TreeDevice() { release(); }
    sayaka_dev_ctx* ctx = nullptr;
    int device = -1;
    unsigned long long topo_version = ~0ull;   // topologyVersion the block tables are for
    unsigned long long surr_version = ~0ull;   // ... and the surrounding tables

   private:
    void release();
};

struct RealBox {
    Vector3d vlo, vhi;
    RealBox() = default;
    RealBox(const Vector3d& lo, const Vector3d& hi) : vlo(lo), vhi(hi) {}
    RealBox(const double lo[], const double hi[])
        : vlo(Vector3d::VecMake(lo[0], lo[1], lo[2])), vhi(Vector3d::VecMake(hi[0], hi[1], hi[2])) {}
    const Vector3d& lo() const { return vlo; }
    const Vector3d& hi() const { return vhi; }
    double center(int d) const { return 0.5 * (vlo(d) + vhi(d)); }
    double length(int d) const { return vhi(d) - vlo(d); }
    Vector3d center() const { return Vector3d::VecMake(center(0), center(1), center(2)); }
    Vector3d length() const { return Vector3d::VecMake(length(0), length(1), length(2)); }
};

struct AmrTreeNode {
    int refineLevel;
    int neighbor[MAX_FACE];
    int child[MAX_CHILD];
    int parent;
    int whichChild;
    int nodeType;
    int nodeType2[MAX_FACE];
    int blockFlags;
    Vector3d blockCenter;
    Vector3d blockLength;
    RealBox boundBox;

    int getLevel() const { return refineLevel; }
    void setLevel(int level) { refineLevel = level; }
    // bounds of a root block set by hand (SayakaTree.h:140-150): box, centre and edge lengths
    void setBoundBox(const RealBox& bb) {
        boundBox = bb;
        blockCenter = bb.center();
        blockLength = bb.length();
    }
    void setBound(const double xlo[], const double xhi[]) { setBoundBox(RealBox(xlo, xhi)); }
    bool isLeaf() const { return nodeType == NodeIsLeaf; }
    bool isNotLeaf() const { return !isLeaf(); }
    bool isParentOfLeafChild() const { return nodeType == NodeHasLeafChild; }
    bool hasChildren() const { return child[0] >= 0; }
    bool hasParent() const { return parent >= 0; }
    void setUnused();
};

// 3^d neighbourhood of a node (SayakaTree.h:336-526): ids, node types and the surrounding
// position each entry was reached through.  Stored [i+1][j+1][k+1] like the reference.
struct SurroundingBlocks {
    int blocks[3][3][3];
    int types[3][3][3];
    int bcrefs[3][3][3];
    int& at(int i, int j, int k) { return blocks[i + 1][j + 1][k + 1]; }
    int at(int i, int j, int k) const { return blocks[i + 1][j + 1][k + 1]; }
    void setUnused();
};

// The blocks of one tree level in ascending id (AmrTreeLevel, SayakaTree.h:154-189).
struct AmrTreeLevel {
    int level = 0;
    int numLevelBlock = 0;
    std::vector<int> levelBlock;
    bool isEmptyLevel() const { return numLevelBlock == 0; }
    const std::vector<int>& blockIds() const { return levelBlock; }
    const int& operator[](int i) const { return levelBlock[i]; }
    int size() const { return numLevelBlock; }
    std::vector<int>::const_iterator begin() const { return levelBlock.begin(); }
    std::vector<int>::const_iterator end() const { return levelBlock.end(); }
    void addBlock(int b) {
        levelBlock.push_back(b);
        numLevelBlock++;
    }
    void clear() {
        levelBlock.clear();
        numLevelBlock = 0;
    }
};

class TreeData;
struct BCPatch;
class FillPatch;

// A state registered with the tree (TreeStateData, SayakaTree.h:528-560): regrid moves its
// blocks along with the tree's, initNewBlockDataAfterRegrid fills the new ones through its
// FillPatch.  The pointers are the caller's, as in the reference.
struct TreeStateData {
    TreeData* old_data = nullptr;
    TreeData* new_data = nullptr;
    BCPatch* bc_patch = nullptr;
    FillPatch* fill_new_patch = nullptr;
    bool hasCurrData() const { return new_data != nullptr; }
    TreeData& currData() { return *new_data; }
    const TreeData& currData() const { return *new_data; }
};

struct AmrTree {
    // configuration (set before init()): either ndim and nb, or the reference's index boxes
    // (blockValidIndex {0,0,0}..{nb-1,nb-1,nb-1}, or ..{nb-1,nb-1,0} in 2D), from which init()
    // takes ndim and nb
    int ndim = 3;
    int nb = 8;            // cells per block edge
    IndexBox blockValidIndex;   // valid cells of a block (SayakaTree.h:207-209)
    IndexBox blockGrownIndex;   // valid and ghost cells
    int blockNumGrow = 1;  // ghost layers of user data
    int minAmrLevel = 1;
    int maxAmrLevel = 1;
    int maxBlockNum = 40000;  // pool size (the reference hard-codes 40000, SayakaTree.cpp:43)

    int maxTreeBlockNum = 0;
    int numBlocks = 0;
    int gridChanged = 0;
    // bumped by every change of the block structure (init, root level, a regrid that changed
    // the tree); the device tables are uploaded again only when it moved
    unsigned long long topologyVersion = 0;

    std::vector<AmrTreeNode> blocks;
    std::vector<int> refineFlag, coarsenFlag, newChildFlag, stayFlag;
    // after regrid(): for every block, the id it had before the regrid, -1 for a block the
    // regrid created (the reference moves the registered state data along the same way,
    // SayakaRegrid.cpp:530-591, and fills the new blocks in initNewBlockDataAfterRegrid)
    std::vector<int> regridOrigin;
    std::vector<SurroundingBlocks> surrBlocks;
    std::vector<AmrTreeLevel> levels;   // levels[l] = the level-l blocks, ascending id
    std::vector<TreeStateData> treeStateData;

    int numFace() const { return 2 * ndim; }
    int numChild() const { return 1 << ndim; }
    int cellsPerBlock() const { return ndim == 3 ? nb * nb * nb : nb * nb; }

    void init();
    void initUniformRootLevel(const int ndiv[3], const RealBox& probbox, const int bc_type[MAX_FACE],
                              const bool is_periodic[MAX_FACE]);
    void initUniformRootLevel(const Vector3i& ndiv, const RealBox& probbox, const int bc_type[MAX_FACE],
                              const bool is_periodic[MAX_FACE]) {
        const int n[3] = {ndiv.x, ndiv.y, ndiv.z};
        initUniformRootLevel(n, probbox, bc_type, is_periodic);
    }
    void initRefineToMinLevel();

    // Returns 0 on success, -1 when the node pool overflows (the reference exits).  The
    // registered states' blocks move with the tree's (SayakaRegrid.cpp:530-591).
    int regrid();
    // AmrTree::initNewBlockDataAfterRegrid (SayakaRegrid.cpp:987-1105) for every registered
    // state, on the GPU through its FillPatch (FillPatch::initNewBlocks).
    void initNewBlockDataAfterRegrid();

    const AmrTreeNode& operator[](int i) const { return blocks[i]; }
    AmrTreeNode& operator[](int i) { return blocks[i]; }

    Vector3d getBlockCellSize(int b) const;
    double getBlockCellVolume(int b) const;
    Vector3d getBlockCellArea(int b) const;
    int currentFinestLevel() const;
    int maxRefineLevel() const { return maxAmrLevel; }
    int minRefineLevel() const { return minAmrLevel; }
    const IndexBox& validBlockCellBox() const { return blockValidIndex; }
    const IndexBox& grownBlockCellBox() const { return blockGrownIndex; }
    const AmrTreeLevel& getTreeLevel(int l) const { return levels[l]; }

    // The tree's device context on GPU `device` (created on first use, sized by maxBlockNum;
    // asking for another GPU than the first throws).
    sayaka_dev_ctx* deviceContext(int device = 0) const;
    // deviceContext with the block tables -- and, with surroundings, the 3^d surrounding
    // tables -- of the current blocks uploaded if the topology changed since their last upload.
    sayaka_dev_ctx* deviceTables(int device = 0, bool surroundings = true) const;
    // the block tables were uploaded by the caller for the current topology (MacSolver's
    // prepareSolver resets the solver's coefficients with them)
    void markTopologyUploaded() const { device_.topo_version = topologyVersion; }

   private:
    mutable TreeDevice device_;
    void checkCoarsen();
    int refineBlocks();
    int coarsenBlocks(int oldBlockNum);
    void setNodeType();
    void buildSurroundingBlocks();
    void cacheTreeLevels();
    int surrIndex(int i, int j, int k) const;
};

}  // namespace amd
}  // namespace sayaka
