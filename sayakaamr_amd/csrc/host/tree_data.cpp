// tree_data.cpp -- host block data (TreeData, SayakaTreeData.h:86-306).
#include "tree_data.h"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

#include "sayaka_amd.h"

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
    const int nz = tree.ndim == 3 ? tree.nb - 1 : 0;
    t.valid_ = IndexBox(Vector3i{0, 0, 0}, Vector3i{tree.nb - 1, tree.nb - 1, nz});
    if (dir >= 0) t.valid_.staggerInDir(dir);
    t.grown_ = t.valid_;
    t.grown_.extend(ngrow, ngrow, t.kg_);
    t.data_.assign(t.blocksize_ * t.nslot_, 0.0);
    t.rebind();
    return t;
}

void DoubleBlockData::setValue(double v) { std::fill(p_, p_ + (std::size_t)ncomp_ * box_.stride(), v); }
void DoubleBlockData::setValue(int scomp, int ncomp, double v) {
    std::fill(p_ + (std::size_t)scomp * box_.stride(), p_ + (std::size_t)(scomp + ncomp) * box_.stride(), v);
}

void TreeData::rebind() {
    views_.assign(nslot_, DoubleBlockData{});
    for (int b = 0; b < nslot_; b++) {
        views_[b].p_ = data_.data() + (std::size_t)b * blocksize_;
        views_[b].box_ = grown_;
        views_[b].ncomp_ = ncomp_;
    }
}

TreeData::TreeData(const TreeData& o)
    : tree_(o.tree_), ndim_(o.ndim_), nb_(o.nb_), ngrow_(o.ngrow_), ncomp_(o.ncomp_), dir_(o.dir_),
      nslot_(o.nslot_), lx_(o.lx_), ly_(o.ly_), lz_(o.lz_), kg_(o.kg_), compsize_(o.compsize_),
      blocksize_(o.blocksize_), valid_(o.valid_), grown_(o.grown_), data_(o.data_) {
    rebind();
}

TreeData::TreeData(TreeData&& o) noexcept { *this = std::move(o); }

TreeData::~TreeData() { releaseDevice(); }

TreeData& TreeData::operator=(const TreeData& o) {
    if (this != &o) {
        TreeData t(o);
        *this = std::move(t);
    }
    return *this;
}

TreeData& TreeData::operator=(TreeData&& o) noexcept {
    if (this == &o) return *this;
    releaseDevice();
    dctx_ = o.dctx_;
    d_data_ = o.d_data_;
    d_tmp_ = o.d_tmp_;
    o.dctx_ = nullptr;
    o.d_data_ = o.d_tmp_ = nullptr;
    tree_ = o.tree_;
    ndim_ = o.ndim_;
    nb_ = o.nb_;
    ngrow_ = o.ngrow_;
    ncomp_ = o.ncomp_;
    dir_ = o.dir_;
    nslot_ = o.nslot_;
    lx_ = o.lx_;
    ly_ = o.ly_;
    lz_ = o.lz_;
    kg_ = o.kg_;
    compsize_ = o.compsize_;
    blocksize_ = o.blocksize_;
    valid_ = o.valid_;
    grown_ = o.grown_;
    data_ = std::move(o.data_);   // the buffer moves: the views keep pointing into it
    views_ = std::move(o.views_);
    o.nslot_ = 0;
    return *this;
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

template <class F>
static void for_blocks(const int iblocks[], int ibegin, int iend, F&& f) {
    for (int q = ibegin; q < iend; q++) f(iblocks[q]);
}

void TreeData::SetValue(TreeData& dst, double val, int dcomp, int ncomp, int ngrow) {
    check_pair(dst, dst, dcomp, dcomp, ncomp, ngrow);
    const int n = dst.tree().numBlocks;
    for (int b = 0; b < n; b++)
        for (int c = 0; c < ncomp; c++)
            for_grown_box(dst, ngrow, [&](int i, int j, int k) { dst(b, i, j, k, dcomp + c) = val; });
}

void TreeData::Copy(TreeData& dst, const TreeData& src, int dcomp, int scomp, int ncomp, int ngrow,
                    const int iblocks[], int ibegin, int iend) {
    check_pair(dst, src, dcomp, scomp, ncomp, ngrow);
    for_blocks(iblocks, ibegin, iend, [&](int b) {
        for (int c = 0; c < ncomp; c++)
            for_grown_box(dst, ngrow, [&](int i, int j, int k) { dst(b, i, j, k, dcomp + c) = src(b, i, j, k, scomp + c); });
    });
}

void TreeData::AddEqual(TreeData& dst, const TreeData& src, double coef, int dcomp, int scomp, int ncomp,
                        int ngrow, const int iblocks[], int ibegin, int iend) {
    check_pair(dst, src, dcomp, scomp, ncomp, ngrow);
    for_blocks(iblocks, ibegin, iend, [&](int b) {
        for (int c = 0; c < ncomp; c++)
            for_grown_box(dst, ngrow, [&](int i, int j, int k) {
                dst(b, i, j, k, dcomp + c) += coef * src(b, i, j, k, scomp + c);
            });
    });
}

void TreeData::SetValue(TreeData& dst, double val, int dcomp, int ncomp, int ngrow, const int iblocks[], int ibegin,
                        int iend) {
    check_pair(dst, dst, dcomp, dcomp, ncomp, ngrow);
    for_blocks(iblocks, ibegin, iend, [&](int b) {
        for (int c = 0; c < ncomp; c++)
            for_grown_box(dst, ngrow, [&](int i, int j, int k) { dst(b, i, j, k, dcomp + c) = val; });
    });
}

void TreeData::setValue(double v) {
    const int n = tree_->numBlocks;
    std::fill(data_.begin(), data_.begin() + (std::size_t)n * blocksize_, v);
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

void TreeData::releaseDevice() {
    if (dctx_) {
        if (d_data_) sayaka_dev_free(dctx_, d_data_);
        if (d_tmp_) sayaka_dev_free(dctx_, d_tmp_);
    }
    dctx_ = nullptr;
    d_data_ = d_tmp_ = nullptr;
}

static void dev_check(int rc, const char* what) {
    if (rc < 0) throw std::runtime_error(std::string("TreeData::") + what + ": " + sayaka_dev_last_error());
}

void TreeData::setResident(bool on, int device) {
    if (!defined()) throw std::logic_error("TreeData::setResident: undefined data");
    if (on == resident()) return;
    if (!on) {
        toHost();
        releaseDevice();
        return;
    }
    sayaka_dev_ctx* c = tree_->deviceContext(device);
    const std::size_t bytes = deviceCompStride() * ncomp_ * sizeof(double);
    double *a = nullptr, *b = nullptr;
    dev_check(sayaka_dev_alloc(c, bytes, (void**)&a), "setResident");
    if (sayaka_dev_alloc(c, bytes, (void**)&b) < 0) {
        const std::string err = sayaka_dev_last_error();
        sayaka_dev_free(c, a);
        throw std::runtime_error("TreeData::setResident: " + err);
    }
    dctx_ = c;
    d_data_ = a;
    d_tmp_ = b;
    toDevice();
}

void TreeData::toDevice() {
    if (!resident()) throw std::logic_error("TreeData::toDevice: the data is not resident");
    const int n = tree_->numBlocks;
    if (n < 1) return;
    stage_.resize((std::size_t)n * compsize_);
    for (int c = 0; c < ncomp_; c++) {
        packGrown(c, n, stage_.data());
        dev_check(sayaka_dev_memcpy(dctx_, d_data_ + c * deviceCompStride(), stage_.data(),
                                    stage_.size() * sizeof(double), 1), "toDevice");
    }
}

void TreeData::toHost() {
    if (!resident()) throw std::logic_error("TreeData::toHost: the data is not resident");
    const int n = tree_->numBlocks;
    if (n < 1) return;
    stage_.resize((std::size_t)n * compsize_);
    for (int c = 0; c < ncomp_; c++) {
        dev_check(sayaka_dev_memcpy(dctx_, stage_.data(), d_data_ + c * deviceCompStride(),
                                    stage_.size() * sizeof(double), 0), "toHost");
        unpackGrown(c, n, stage_.data());
    }
}

void TreeData::swapDeviceBuffers() { std::swap(d_data_, d_tmp_); }

void TreeData::remapBlocks(const std::vector<int>& origin, int nblk) {
    if (resident()) {
        // the device copy is the data: move its blocks there (the host copy is left as it was)
        if ((int)origin.size() < nblk) throw std::invalid_argument("TreeData::remapBlocks: origin too short");
        dev_check(sayaka_dev_block_move(dctx_, d_data_, d_tmp_, ncomp_, (long long)deviceCompStride(),
                                        (long long)compsize_, nblk, nslot_, origin.data()),
                  "remapBlocks");
        swapDeviceBuffers();
        return;
    }
    std::vector<double> out(data_.size(), 0.0);
    for (int b = 0; b < nblk && b < nslot_; b++)
        if (origin[b] >= 0)
            std::memcpy(out.data() + (std::size_t)b * blocksize_, data_.data() + (std::size_t)origin[b] * blocksize_,
                        blocksize_ * sizeof(double));
    data_.swap(out);
    rebind();
}

}  // namespace amd
}  // namespace sayaka
