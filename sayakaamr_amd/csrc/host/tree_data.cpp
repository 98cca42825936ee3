// tree_data.cpp -- host block data (TreeData, SayakaTreeData.h:86-306).
#include "tree_data.h"

#include <algorithm>
#include <cstring>
#include <stdexcept>

namespace sayaka {
inline namespace amd {

TreeData TreeData::MakeCellData(const AmrTree& tree, int ncomp, int ngrow) {
    return MakeFaceData(tree, -1, ncomp, ngrow);
}

TreeData TreeData::MakeFaceData(const AmrTree& tree, int dir, int ncomp, int ngrow) {
    if (ncomp < 1 || ngrow < 0) throw std::invalid_argument("TreeData: ncomp >= 1 and ngrow >= 0 required");
    if (dir >= tree.ndim) throw std::invalid_argument("TreeData: face direction out of range");
    TreeData t;
    t.tree_ = &tree;
    t.ndim_ = tree.ndim;
    t.nb_ = tree.nb;
    t.ngrow_ = ngrow;
    t.ncomp_ = ncomp;
    t.dir_ = dir;
    t.nslot_ = tree.maxBlockNum;
    t.kg_ = tree.ndim == 3 ? ngrow : 0;
    t.lx_ = tree.nb + 2 * ngrow + (dir == 0);
    t.ly_ = tree.nb + 2 * ngrow + (dir == 1);
    t.lz_ = tree.ndim == 3 ? tree.nb + 2 * ngrow + (dir == 2) : 1;
    t.compsize_ = (std::size_t)t.lx_ * t.ly_ * t.lz_;
    t.blocksize_ = t.compsize_ * ncomp;
    t.data_.assign(t.blocksize_ * t.nslot_, 0.0);
    return t;
}

TreeData* TreeData::CreateCellData(const AmrTree& tree, int ncomp, int ngrow) {
    return new TreeData(MakeCellData(tree, ncomp, ngrow));
}

TreeData* TreeData::CreateFaceData(const AmrTree& tree, int dir, int ncomp, int ngrow) {
    return new TreeData(MakeFaceData(tree, dir, ncomp, ngrow));
}

std::vector<TreeData*> TreeData::CreateFaceDataPArray(const AmrTree& tree, int ncomp, int ngrow) {
    std::vector<TreeData*> out;
    for (int d = 0; d < tree.ndim; d++) out.push_back(CreateFaceData(tree, d, ncomp, ngrow));
    return out;
}

void TreeData::ReleaseDataPArray(std::vector<TreeData*>& parray, bool clear_array) {
    for (TreeData*& p : parray) {
        delete p;
        p = nullptr;
    }
    if (clear_array) parray.clear();
}

template <class F>
static void for_grown_box(const TreeData& d, int ngrow, F&& f) {
    const int gz = d.extent(2) > 1 ? ngrow : 0;
    for (int k = -gz; k < d.validExtent(2) + gz; k++)
        for (int j = -ngrow; j < d.validExtent(1) + ngrow; j++)
            for (int i = -ngrow; i < d.validExtent(0) + ngrow; i++) f(i, j, k);
}

static void check_pair(const TreeData& dst, const TreeData& src, int dcomp, int scomp, int ncomp, int ngrow) {
    if (dst.faceDir() != src.faceDir() || dcomp < 0 || scomp < 0 || dcomp + ncomp > dst.numComp() ||
        scomp + ncomp > src.numComp() || ngrow < 0 || ngrow > dst.numGrow() || ngrow > src.numGrow())
        throw std::invalid_argument("TreeData: incompatible data, components or ghost layers");
}

void TreeData::Copy(TreeData& dst, const TreeData& src, int dcomp, int scomp, int ncomp, int ngrow) {
    check_pair(dst, src, dcomp, scomp, ncomp, ngrow);
    const int n = dst.tree().numBlocks;
    for (int b = 0; b < n; b++)
        for (int c = 0; c < ncomp; c++)
            for_grown_box(dst, ngrow, [&](int i, int j, int k) { dst(b, i, j, k, dcomp + c) = src(b, i, j, k, scomp + c); });
}

void TreeData::AddEqual(TreeData& dst, const TreeData& src, double coef, int dcomp, int scomp, int ncomp, int ngrow) {
    check_pair(dst, src, dcomp, scomp, ncomp, ngrow);
    const int n = dst.tree().numBlocks;
    for (int b = 0; b < n; b++)
        for (int c = 0; c < ncomp; c++)
            for_grown_box(dst, ngrow, [&](int i, int j, int k) {
                dst(b, i, j, k, dcomp + c) += coef * src(b, i, j, k, scomp + c);
            });
}

int TreeData::validExtent(int d) const {
    if (d == 2 && ndim_ == 2) return 1;
    return nb_ + (d == dir_ ? 1 : 0);
}

void TreeData::SetValue(double v) { std::fill(data_.begin(), data_.end(), v); }

void TreeData::SetValue(double v, int comp) {
    for (int b = 0; b < nslot_; b++)
        std::fill(data_.begin() + (std::size_t)b * blocksize_ + (std::size_t)comp * compsize_,
                  data_.begin() + (std::size_t)b * blocksize_ + (std::size_t)(comp + 1) * compsize_, v);
}

void TreeData::packValid(int comp, int nblk, double* out) const {
    const int ex = validExtent(0), ey = validExtent(1), ez = validExtent(2);
    std::size_t q = 0;
    for (int b = 0; b < nblk; b++)
        for (int k = 0; k < ez; k++)
            for (int j = 0; j < ey; j++)
                for (int i = 0; i < ex; i++) out[q++] = (*this)(b, i, j, k, comp);
}

void TreeData::unpackValid(int comp, int nblk, const double* in) {
    const int ex = validExtent(0), ey = validExtent(1), ez = validExtent(2);
    std::size_t q = 0;
    for (int b = 0; b < nblk; b++)
        for (int k = 0; k < ez; k++)
            for (int j = 0; j < ey; j++)
                for (int i = 0; i < ex; i++) (*this)(b, i, j, k, comp) = in[q++];
}

void TreeData::packGrown(int comp, int nblk, double* out) const {
    for (int b = 0; b < nblk; b++)
        std::memcpy(out + (std::size_t)b * compsize_,
                    data_.data() + (std::size_t)b * blocksize_ + (std::size_t)comp * compsize_,
                    compsize_ * sizeof(double));
}

void TreeData::unpackGrown(int comp, int nblk, const double* in) {
    for (int b = 0; b < nblk; b++)
        std::memcpy(data_.data() + (std::size_t)b * blocksize_ + (std::size_t)comp * compsize_,
                    in + (std::size_t)b * compsize_, compsize_ * sizeof(double));
}

void TreeData::remapBlocks(const std::vector<int>& origin, int nblk) {
    std::vector<double> out(data_.size(), 0.0);
    for (int b = 0; b < nblk && b < nslot_; b++)
        if (origin[b] >= 0)
            std::memcpy(out.data() + (std::size_t)b * blocksize_, data_.data() + (std::size_t)origin[b] * blocksize_,
                        blocksize_ * sizeof(double));
    data_.swap(out);
}

}  // namespace amd
}  // namespace sayaka
