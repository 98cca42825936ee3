// fill_patch.cpp -- FillPatch on the GPU (see fill_patch.h).
#include "fill_patch.h"

#include <cstddef>
#include <set>
#include <stdexcept>
#include <string>

#include "dev_tree.h"
#include "sayaka_amd.h"

namespace sayaka {
inline namespace amd {

namespace {
void check(int rc, const char* what) {
    if (rc < 0) throw std::runtime_error(std::string(what) + ": " + sayaka_dev_last_error());
}
}  // namespace

FillPatch::FillPatch(const AmrTree& tree, TreeData& data, int device)
    : tree_(&tree), data_(&data), prolong_(data.numComp(), InterpPatch::PROLONG_INJECTION),
      restrict_(data.numComp(), InterpPatch::RESTRICT_AVERAGE), device_(device) {}

FillPatch::~FillPatch() {
    // the context is the tree's (it outlives the patch); the staging buffers are the patch's
    if (ctx_) {
        if (d_buf_[0]) sayaka_dev_free(ctx_, d_buf_[0]);
        if (d_buf_[1]) sayaka_dev_free(ctx_, d_buf_[1]);
    }
}

void FillPatch::setProlongation(int comp, int prolong) { prolong_.at(comp) = prolong; }

void FillPatch::prepareDevice() {
    // the tree's context; its tables go up only when the tree changed since the last upload
    ctx_ = tree_->deviceTables(device_, true);
}

double* FillPatch::stagingBuffer(int which, std::size_t bytes) {
    // grown only, kept between calls: no allocation per fill in steady state
    if (bytes > buf_bytes_[which]) {
        if (d_buf_[which]) sayaka_dev_free(ctx_, d_buf_[which]);
        d_buf_[which] = nullptr;
        buf_bytes_[which] = 0;
        check(sayaka_dev_alloc(ctx_, bytes, (void**)&d_buf_[which]), "FillPatch: device state");
        buf_bytes_[which] = bytes;
    }
    return d_buf_[which];
}

void FillPatch::bcTables(int scomp, int ncomp, std::vector<int>& codes, std::vector<int>& mathbc,
                         std::vector<double>& bcval) const {
    // BC register of the filled components in the C ABI's layout
    codes.clear();
    mathbc.clear();
    bcval.clear();
    if (data_->faceDir() >= 0 || !bc_) return;
    std::set<int> cs;
    for (int c = 0; c < ncomp; c++)
        for (const auto& m : bc_->boundaryRegister(scomp + c).bcmaps)
            for (const auto& kv : m) cs.insert(kv.first);
    codes.assign(cs.begin(), cs.end());
    const int nphys = (int)codes.size();
    mathbc.assign((size_t)ncomp * 27 * nphys, -1);
    bcval.assign((size_t)ncomp * 27 * nphys, 0.0);
    for (int c = 0; c < ncomp; c++) {
        const BCRegister& r = bc_->boundaryRegister(scomp + c);
        for (int ii = -1; ii <= 1; ii++)
            for (int jj = -1; jj <= 1; jj++)
                for (int kk = -1; kk <= 1; kk++) {
                    const int q = (ii + 1) * 9 + (jj + 1) * 3 + (kk + 1);
                    for (int h = 0; h < nphys; h++) {
                        mathbc[((size_t)c * 27 + q) * nphys + h] = r.getBCMap(ii, jj, kk, codes[h]);
                        bcval[((size_t)c * 27 + q) * nphys + h] = r.getBCVal(ii, jj, kk, codes[h]);
                    }
                }
    }
}

void FillPatch::initNewBlocks(const std::vector<int>& isnew) {
    const int nblk = tree_->numBlocks, ncomp = data_->numComp(), fdir = data_->faceDir();
    for (int p : prolong_)
        if (p == InterpPatch::PROLONG_IGNORE)
            throw std::invalid_argument("FillPatch: PROLONG_IGNORE components are not filled by this build");
    std::vector<int> codes, mathbc;
    std::vector<double> bcval;
    bcTables(0, ncomp, codes, mathbc, bcval);
    std::vector<int> origin(nblk);
    for (int b = 0; b < nblk; b++) origin[b] = isnew[b] ? -1 : b;   // the blocks have moved already
    prepareDevice();
    if (data_->resident()) {
        // the device copy is the data: fill the new blocks there, no host transfer
        const int nslot = data_->numBlockSlots();
        check(sayaka_dev_state_remap(ctx_, data_->deviceData(), nslot, nblk, data_->deviceScratch(), ncomp, nslot,
                                     data_->numGrow(), fdir, origin.data(), prolong_.data(), (int)codes.size(),
                                     codes.data(), mathbc.data(), bcval.data()),
              "FillPatch::initNewBlocks");
        data_->swapDeviceBuffers();
        return;
    }
    const size_t per = (size_t)nblk * data_->grownSize();
    host_.resize(per * ncomp);
    for (int c = 0; c < ncomp; c++) data_->packGrown(c, nblk, host_.data() + c * per);
    double* d_old = stagingBuffer(0, host_.size() * sizeof(double));
    double* d_new = stagingBuffer(1, host_.size() * sizeof(double));
    check(sayaka_dev_memcpy(ctx_, d_old, host_.data(), host_.size() * sizeof(double), 1), "FillPatch::initNewBlocks");
    check(sayaka_dev_state_remap(ctx_, d_old, nblk, nblk, d_new, ncomp, nblk, data_->numGrow(), fdir, origin.data(),
                                 prolong_.data(), (int)codes.size(), codes.data(), mathbc.data(), bcval.data()),
          "FillPatch::initNewBlocks");
    check(sayaka_dev_memcpy(ctx_, host_.data(), d_new, host_.size() * sizeof(double), 0), "FillPatch::initNewBlocks");
    for (int c = 0; c < ncomp; c++) data_->unpackGrown(c, nblk, host_.data() + c * per);
}

void FillPatch::fillBoundary(int scomp, int ncomp, int nlayer) {
    if (nlayer <= 0 || ncomp <= 0) return;   // SayakaFillPatch.cpp:21
    if (scomp < 0 || scomp + ncomp > data_->numComp()) throw std::invalid_argument("FillPatch: component range");
    const int fdir = data_->faceDir();
    const int nblk = tree_->numBlocks;
    std::vector<int> prolong(prolong_.begin() + scomp, prolong_.begin() + scomp + ncomp);
    for (int p : prolong)
        if (p == InterpPatch::PROLONG_IGNORE)
            throw std::invalid_argument("FillPatch: PROLONG_IGNORE components are not filled by this build");
    std::vector<int> codes, mathbc;
    std::vector<double> bcval;
    bcTables(scomp, ncomp, codes, mathbc, bcval);
    prepareDevice();
    auto fill = [&](double* state, int nc_total, int blocks_alloc, int sc) {
        if (fdir < 0)
            return sayaka_dev_state_fill(ctx_, state, nc_total, blocks_alloc, data_->numGrow(), sc, ncomp, nlayer,
                                         prolong.data(), (int)codes.size(), codes.data(), mathbc.data(), bcval.data());
        return sayaka_dev_face_fill(ctx_, state, nc_total, blocks_alloc, fdir, data_->numGrow(), sc, ncomp, nlayer,
                                    prolong.data());
    };
    if (data_->resident()) {
        // the device copy is the data: fill it in place, no host transfer
        check(fill(data_->deviceData(), data_->numComp(), data_->numBlockSlots(), scomp), "FillPatch::fillBoundary");
        return;
    }
    const size_t per = (size_t)nblk * data_->grownSize();
    host_.resize(per * ncomp);
    for (int c = 0; c < ncomp; c++) data_->packGrown(scomp + c, nblk, host_.data() + c * per);
    double* d = stagingBuffer(0, host_.size() * sizeof(double));
    check(sayaka_dev_memcpy(ctx_, d, host_.data(), host_.size() * sizeof(double), 1), "FillPatch::fillBoundary");
    check(fill(d, ncomp, nblk, 0), "FillPatch::fillBoundary");
    check(sayaka_dev_memcpy(ctx_, host_.data(), d, host_.size() * sizeof(double), 0), "FillPatch::fillBoundary");
    for (int c = 0; c < ncomp; c++) data_->unpackGrown(scomp + c, nblk, host_.data() + c * per);
}

}  // namespace amd
}  // namespace sayaka
