// fill_patch.h -- ghost fill of a caller's TreeData on the GPU (the reference's user fills).
//
// Mirrors InterpPatch::ProlongMethod (SayakaInterpPatch.h:78-83), BCPatch
// (SayakaBoundaryPatch.h:177-260: one BCRegister per component) and FillPatch
// (SayakaFillPatch.h:60-200: setBC, setProlongation, fillBoundary), so that a driver's
//     FillPatch fp(tree, data); fp.setBC(bc); fp.setProlongation(c, m); fp.fillBoundary(0, n, 2);
// compiles unchanged.  fillBoundary runs on the device (sayaka_dev_state_fill for cell data,
// sayaka_dev_face_fill for face data) in the tree's device context, shared by every patch and
// solver of the tree; the tree's tables go up only when the tree changed since their last
// upload, and the device staging buffers are kept between calls.  Data kept on the device
// (TreeData::setResident) is filled in place with no host transfer; other data's grown boxes
// go up and the filled ghosts come back.  The result is the reference's
// FillPatch::fillBoundary (SayakaFillPatch.cpp:17-368) bit for bit.  Errors throw
// std::runtime_error where the reference exits (an unmapped wall code, face data next to a wall).
#pragma once

#include <cstddef>
#include <vector>

#include "bc_register.h"
#include "tree_data.h"

struct sayaka_dev_ctx;

namespace sayaka {
inline namespace amd {

struct InterpPatch {
    enum ProlongMethod {
        PROLONG_IGNORE = -1,
        PROLONG_INJECTION = 0,
        PROLONG_CENTERED_LINEAR,
        PROLONG_CENTERED_LINEAR_LIMITED,
    };
    // restriction methods (SayakaInterpPatch.h:85-92); recorded by FillPatch::setRestriction
    enum RestrictMethod { RESTRICT_IGNORE = -1, RESTRICT_AVERAGE = 0, RESTRICT_SUM };
};

struct BCPatch {
    std::vector<BCRegister> regs;
    BCPatch() = default;
    // BCPatch(tree, data, location) / define(tree, location, ncomp, ngrow)
    // (SayakaBoundaryPatch.h:196-215): one BC register per component
    BCPatch(const AmrTree&, const TreeData& data, VariableLocation) { defineByData(data); }
    void define(const AmrTree&, VariableLocation, int ncomp, int) { regs.assign(ncomp, BCRegister{}); }
    void defineByData(const TreeData& data) { regs.assign(data.numComp(), BCRegister{}); }
    BCRegister& boundaryRegister(int comp) { return regs.at(comp); }
    const BCRegister& boundaryRegister(int comp) const { return regs.at(comp); }
};

class FillPatch {
   public:
    // The tree and the data must outlive the FillPatch.  The fills run in the tree's device
    // context on the given GPU (AmrTree::deviceContext).
    FillPatch(const AmrTree& tree, TreeData& data, int device = 0);
    ~FillPatch();
    FillPatch(const FillPatch&) = delete;
    FillPatch& operator=(const FillPatch&) = delete;

    void setBC(const BCPatch& bc) { bc_ = &bc; }
    // default: injection for every component (SayakaInterpPatch.h:109)
    void setProlongation(int comp, int prolong);
    // restriction method per component (InterpPatch::setRestriction, SayakaInterpPatch.h:119-122;
    // default average); kept with the patch -- no fill here restricts
    void setRestriction(int comp, int method) { restrict_.at(comp) = method; }
    const std::vector<int>& prolongFlag() const { return prolong_; }
    const std::vector<int>& restrictFlag() const { return restrict_; }
    // FillPatch::fillBoundary (SayakaFillPatch.cpp:17-30): levels 1..finest, every block
    void fillBoundary(int scomp, int ncomp, int nlayer);
    // initNewBlockDataAfterRegrid's work on this state (SayakaRegrid.cpp:1013-1103), after the
    // regrid moved its blocks: the parents of new blocks filled in-level, the new blocks
    // (isnew[b] != 0) prolonged, face data corrected from existing neighbours; on the GPU
    // (sayaka_dev_state_remap).  Called by AmrTree::initNewBlockDataAfterRegrid.
    void initNewBlocks(const std::vector<int>& isnew);

   private:
    void prepareDevice();
    double* stagingBuffer(int which, std::size_t bytes);
    void bcTables(int scomp, int ncomp, std::vector<int>& codes, std::vector<int>& mathbc,
                  std::vector<double>& bcval) const;
    const AmrTree* tree_;
    TreeData* data_;
    const BCPatch* bc_ = nullptr;
    std::vector<int> prolong_, restrict_;
    int device_;
    sayaka_dev_ctx* ctx_ = nullptr;         // the tree's
    double* d_buf_[2] = {nullptr, nullptr};  // device staging of host data, kept between calls
    std::size_t buf_bytes_[2] = {0, 0};
    std::vector<double> host_;               // host staging, kept between calls
};

}  // namespace amd
}  // namespace sayaka
