// tree_data.h -- block data of one tree on the host (the caller's fields).
//
// Mirrors the reference TreeData / GridData (SayakaTreeData.h:86-306, SayakaBlockData.h:12-215):
// one slot per node of the tree's pool, each block stored [comp][k][j][i] over its box grown
// by ngrow cells (i fastest, SayakaBox.h:214-217); face data is staggered in one direction
// (one more face than cells along it).  Valid cells are indexed from 0, ghosts from -ngrow.
// The solver on the device keeps its own packed valid-only copies; these host objects are
// what callers fill and read, exactly as with the reference.  A TreeData can also keep a copy
// on the device (setResident): then the tree's ghost fills, regrid moves and new-block fills
// run on that copy with no host transfer, and the host copy is synchronised on request.
#pragma once

#include <cstddef>
#include <vector>

#include "amr_tree.h"

struct sayaka_dev_ctx;

namespace sayaka {
inline namespace amd {

// The data of one block, over its grown box (GridData<double>, SayakaBlockData.h:12-215):
// indexed (i, j, k, comp) with valid cells from 0 and ghosts from -ngrow.  TreeData owns the
// storage; TreeData::operator[] hands out references to these views.
class DoubleBlockData {
   public:
    double& operator()(int i, int j, int k, int comp = 0) { return p_[box_.offset(i, j, k, comp)]; }
    double operator()(int i, int j, int k, int comp = 0) const { return p_[box_.offset(i, j, k, comp)]; }
    const IndexBox& box() const { return box_; }
    int numComp() const { return ncomp_; }
    // every value of the grown box, every component / components [scomp, scomp + ncomp)
    void setValue(double v);
    void setValue(int scomp, int ncomp, double v);
    double* dataPtr(int comp = 0) { return p_ + (std::size_t)comp * box_.stride(); }
    const double* dataPtr(int comp = 0) const { return p_ + (std::size_t)comp * box_.stride(); }

   private:
    friend class TreeData;
    double* p_ = nullptr;
    IndexBox box_;
    int ncomp_ = 0;
};
using GridDataD = DoubleBlockData;

class TreeData {
   public:
    TreeData() = default;
    ~TreeData();
    TreeData(const TreeData& o);   // copies the host data; the copy is not resident
    TreeData(TreeData&& o) noexcept;
    TreeData& operator=(const TreeData& o);
    TreeData& operator=(TreeData&& o) noexcept;

    // TreeData::MakeCellData / MakeFaceData (SayakaTreeData.h:245-306)
    static TreeData MakeCellData(const AmrTree& tree, int ncomp, int ngrow);
    static TreeData MakeFaceData(const AmrTree& tree, int dir, int ncomp, int ngrow);
    // heap-allocated forms (SayakaTreeData.h:227-232); the caller deletes them
    static TreeData* CreateCellData(const AmrTree& tree, int ncomp, int ngrow);
    static TreeData* CreateFaceData(const AmrTree& tree, int dir, int ncomp, int ngrow);
    static std::vector<TreeData*> CreateFaceDataPArray(const AmrTree& tree, int ncomp, int ngrow);
    static void ReleaseDataPArray(std::vector<TreeData*>& parray, bool clear_array = true);
    // TreeData::Copy / AddEqual (SayakaTreeData.cpp:230-340): components [scomp, scomp+ncomp) of
    // src into [dcomp, ...) of dst over the valid box grown by ngrow, every block of the tree
    // (AddEqual: dst += coef * src)
    static void Copy(TreeData& dst, const TreeData& src, int dcomp, int scomp, int ncomp, int ngrow);
    static void AddEqual(TreeData& dst, const TreeData& src, double coef, int dcomp, int scomp, int ncomp,
                         int ngrow);
    static void SetValue(TreeData& dst, double val, int dcomp, int ncomp, int ngrow);
    // the same over a list of blocks, iblocks[ibegin .. iend) (SayakaTreeData.h:256-294; e.g.
    // a tower's blockIds())
    static void Copy(TreeData& dst, const TreeData& src, int dcomp, int scomp, int ncomp, int ngrow,
                     const int iblocks[], int ibegin, int iend);
    static void Copy(TreeData& dst, const TreeData& src, int dcomp, int scomp, int ncomp, int ngrow,
                     const std::vector<int>& iblocks) {
        Copy(dst, src, dcomp, scomp, ncomp, ngrow, iblocks.data(), 0, (int)iblocks.size());
    }
    static void AddEqual(TreeData& dst, const TreeData& src, double coef, int dcomp, int scomp, int ncomp,
                         int ngrow, const int iblocks[], int ibegin, int iend);
    static void AddEqual(TreeData& dst, const TreeData& src, double coef, int dcomp, int scomp, int ncomp,
                         int ngrow, const std::vector<int>& iblocks) {
        AddEqual(dst, src, coef, dcomp, scomp, ncomp, ngrow, iblocks.data(), 0, (int)iblocks.size());
    }
    static void SetValue(TreeData& dst, double val, int dcomp, int ncomp, int ngrow, const int iblocks[],
                         int ibegin, int iend);
    static void SetValue(TreeData& dst, double val, int dcomp, int ncomp, int ngrow,
                         const std::vector<int>& iblocks) {
        SetValue(dst, val, dcomp, ncomp, ngrow, iblocks.data(), 0, (int)iblocks.size());
    }

    // block views (SayakaTreeData.h:176-177)
    DoubleBlockData& operator[](int iblock) { return views_[iblock]; }
    const DoubleBlockData& operator[](int iblock) const { return views_[iblock]; }
    // valid box (cells, or faces staggered in the face direction) and grown box of a block
    const IndexBox& validBox() const { return valid_; }
    const IndexBox& indexBox() const { return grown_; }
    VariableLocation indexType() const { return valid_.type(); }
    int isCellData() const { return dir_ < 0; }
    int isFaceData() const { return dir_ >= 0; }
    int isFaceData(int d) const { return dir_ == d; }
    int getFaceDataDir() const { return dir_; }
    int blockNum() const { return tree_->numBlocks; }
    int maxBlockNum() const { return nslot_; }
    // every value of the tree's current blocks (TreeData::setValue, SayakaTreeData.h:187-192)
    void setValue(double v);

    // the tree the data was made for (TreeData::tree, SayakaTreeData.h:149); it must outlive the data
    const AmrTree& tree() const { return *tree_; }
    int numComp() const { return ncomp_; }
    int numGrow() const { return ngrow_; }
    int faceDir() const { return dir_; }   // -1 for cell data
    int numBlockSlots() const { return nslot_; }
    bool defined() const { return nslot_ > 0; }

    // extents of the grown box along each direction
    int extent(int d) const { return d == 0 ? lx_ : (d == 1 ? ly_ : lz_); }
    // valid extent along d (cells, or faces in the staggered direction)
    int validExtent(int d) const;

    double& operator()(int blk, int i, int j, int k, int comp) { return data_[offset(blk, i, j, k, comp)]; }
    double operator()(int blk, int i, int j, int k, int comp) const { return data_[offset(blk, i, j, k, comp)]; }

    // TreeData::SetValue (SayakaTreeData.cpp:230-263): every slot, grown region included
    void SetValue(double v);
    void SetValue(double v, int comp);

    // Valid region of component comp of blocks [0, nblk) to / from a packed array
    // [blk][k][j][i] (the layout of the device C ABI).
    void packValid(int comp, int nblk, double* out) const;
    void unpackValid(int comp, int nblk, const double* in);
    // Component comp of blocks [0, nblk), grown region included, to a packed array
    // [blk][grown box] (the layout sayaka_dev_fix_bc reads).
    void packGrown(int comp, int nblk, double* out) const;
    void unpackGrown(int comp, int nblk, const double* in);
    // the regrid's packing: slot b takes slot origin[b] (the whole block), 0 where origin is
    // -1; slots from nblk on are zeroed (SayakaRegrid.cpp:530-591)
    void remapBlocks(const std::vector<int>& origin, int nblk);
    std::size_t grownSize() const { return compsize_; }   // values per block and component

    // ---- device copy (no reference counterpart: the reference's data lives in host memory)
    // setResident(true) makes a copy in the tree's device context (AmrTree::deviceContext) in
    // the C ABI's layout [comp][slot][grown box] and uploads the host data of the tree's
    // current blocks.  From then on FillPatch::fillBoundary, AmrTree::regrid's block moves and
    // initNewBlockDataAfterRegrid work on the device copy only, with no host transfer; the host
    // copy is brought up to date by toHost(), and host changes go to the device by toDevice().
    // setResident(false) brings the data back to the host and frees the device copy.  The
    // tree must outlive the data.
    void setResident(bool on, int device = 0);
    bool resident() const { return d_data_ != nullptr; }
    void toDevice();
    void toHost();
    double* deviceData() const { return d_data_; }
    // values between two components of the device copy (slots x grown box)
    std::size_t deviceCompStride() const { return (std::size_t)nslot_ * compsize_; }
    // the device copy's second buffer, which a block move or a new-block fill writes before
    // swapDeviceBuffers() makes it the data
    double* deviceScratch() const { return d_tmp_; }
    void swapDeviceBuffers();
    sayaka_dev_ctx* deviceContext() const { return dctx_; }

   private:
    void releaseDevice();
    std::size_t offset(int blk, int i, int j, int k, int comp) const {
        return (std::size_t)blk * blocksize_ + (std::size_t)comp * compsize_ +
               (std::size_t)((i + ngrow_) + lx_ * ((j + ngrow_) + ly_ * (k + kg_)));
    }
    void rebind();   // point the block views at data_
    const AmrTree* tree_ = nullptr;
    int ndim_ = 3, nb_ = 8, ngrow_ = 0, ncomp_ = 0, dir_ = -1, nslot_ = 0;
    int lx_ = 0, ly_ = 0, lz_ = 0, kg_ = 0;
    std::size_t compsize_ = 0, blocksize_ = 0;
    IndexBox valid_, grown_;
    std::vector<double> data_;
    std::vector<DoubleBlockData> views_;
    sayaka_dev_ctx* dctx_ = nullptr;   // the tree's context, while resident
    double* d_data_ = nullptr;
    double* d_tmp_ = nullptr;
    std::vector<double> stage_;        // host staging of one component in the device layout
};

}  // namespace amd
}  // namespace sayaka
