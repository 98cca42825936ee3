// fill_patch.cpp -- FillPatch on the GPU (see fill_patch.h).
#include "fill_patch.h"

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
    if (ctx_) sayaka_dev_destroy(ctx_);
}

void FillPatch::setProlongation(int comp, int prolong) { prolong_.at(comp) = prolong; }

void FillPatch::prepareDevice() {
    if (!ctx_)
        check(sayaka_dev_create(device_, tree_->ndim, tree_->nb, tree_->maxBlockNum, &ctx_), "FillPatch: device context");
    // the tree may have been regridded since the last fill: tables every time
    upload_tree_topology(ctx_, *tree_);
    upload_tree_surroundings(ctx_, *tree_);
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
    const size_t per = (size_t)nblk * data_->grownSize();
    std::vector<double> host(per * ncomp);
    for (int c = 0; c < ncomp; c++) data_->packGrown(c, nblk, host.data() + c * per);
    double *d_old = nullptr, *d_new = nullptr;
    check(sayaka_dev_alloc(ctx_, host.size() * sizeof(double), (void**)&d_old), "FillPatch: device state");
    int rc = sayaka_dev_alloc(ctx_, host.size() * sizeof(double), (void**)&d_new);
    if (rc >= 0) rc = sayaka_dev_memcpy(ctx_, d_old, host.data(), host.size() * sizeof(double), 1);
    if (rc >= 0)
        rc = sayaka_dev_state_remap(ctx_, d_old, nblk, nblk, d_new, ncomp, nblk, data_->numGrow(), fdir, origin.data(),
                                    prolong_.data(), (int)codes.size(), codes.data(), mathbc.data(), bcval.data());
    if (rc >= 0) rc = sayaka_dev_memcpy(ctx_, host.data(), d_new, host.size() * sizeof(double), 0);
    const std::string err = rc < 0 ? sayaka_dev_last_error() : "";
    sayaka_dev_free(ctx_, d_old);
    if (d_new) sayaka_dev_free(ctx_, d_new);
    if (rc < 0) throw std::runtime_error("FillPatch::initNewBlocks: " + err);
    for (int c = 0; c < ncomp; c++) data_->unpackGrown(c, nblk, host.data() + c * per);
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
    const size_t per = (size_t)nblk * data_->grownSize();
    std::vector<double> host(per * ncomp);
    for (int c = 0; c < ncomp; c++) data_->packGrown(scomp + c, nblk, host.data() + c * per);
    double* d = nullptr;
    check(sayaka_dev_alloc(ctx_, host.size() * sizeof(double), (void**)&d), "FillPatch: device state");
    int rc = sayaka_dev_memcpy(ctx_, d, host.data(), host.size() * sizeof(double), 1);
    if (rc >= 0) {
        if (fdir < 0)
            rc = sayaka_dev_state_fill(ctx_, d, ncomp, nblk, data_->numGrow(), 0, ncomp, nlayer, prolong.data(),
                                       (int)codes.size(), codes.data(), mathbc.data(), bcval.data());
        else
            rc = sayaka_dev_face_fill(ctx_, d, ncomp, nblk, fdir, data_->numGrow(), 0, ncomp, nlayer, prolong.data());
    }
    if (rc >= 0) rc = sayaka_dev_memcpy(ctx_, host.data(), d, host.size() * sizeof(double), 0);
    const std::string err = rc < 0 ? sayaka_dev_last_error() : "";
    sayaka_dev_free(ctx_, d);
    if (rc < 0) throw std::runtime_error("FillPatch::fillBoundary: " + err);
    for (int c = 0; c < ncomp; c++) data_->unpackGrown(scomp + c, nblk, host.data() + c * per);
}

}  // namespace amd
}  // namespace sayaka
