// index_box.h -- index boxes, variable locations and surrounding / face indices.
//
// The value types a driver of the reference writes its block loops with (SayakaBox.h:24-410,
// SayakaSurroundingIndex.h:13-120, SayakaFaceIndex.h:10-53, SayakaCommons.h:40-70), with the
// same member names so that such loops compile unchanged:
//     IndexBox vb = tree.validBlockCellBox();
//     BEGIN_FOR_BOX_RANGE(vb, i, j, k); { d(i, j, k, 0) = ...; } END_FOR_BOX_RANGE;
//
// The library itself is dimension-agnostic (a tree carries its dimension).  A driver is
// compiled for one dimension, as the reference is: NDIM below is SAYAKA_SPACEDIM (default 3)
// and only the driver-facing helpers that the reference ties to NDIM use it (IndexBox::extend
// with one argument, SurroundingIndex, FaceIndex); nothing inside the library calls them.
#pragma once

#include <cassert>
#include <tuple>

#include "sayaka_types.h"

#ifndef SAYAKA_SPACEDIM
#define SAYAKA_SPACEDIM 3
#endif

namespace sayaka {
inline namespace amd {

constexpr int NDIM = SAYAKA_SPACEDIM;
static_assert(NDIM == 2 || NDIM == 3, "SAYAKA_SPACEDIM must be 2 or 3");

// Where a variable sits in a cell: bit d set = staggered (on faces) in direction d.
enum { VARLOC_CELL = 0, VARLOC_FACE_X = 1, VARLOC_FACE_Y = 2, VARLOC_FACE_Z = 4 };

struct VariableLocation {
    int bits = VARLOC_CELL;
    constexpr VariableLocation() = default;
    constexpr VariableLocation(int b) : bits(b) {}
    int isStaggered(int d) const { return (bits >> d) & 1; }
    void setStaggered(int d) { bits |= 1 << d; }
    int isCellVar() const { return bits == VARLOC_CELL; }
    int isFaceVar(int d) const { return bits == (1 << d); }
    int isFaceVar() const { return bits == VARLOC_FACE_X || bits == VARLOC_FACE_Y || bits == VARLOC_FACE_Z; }
    int getFaceVarDir() const {
        for (int d = 0; d < MAX_DIM; d++)
            if (isFaceVar(d)) return d;
        return -1;
    }
    operator int() const { return bits; }
};

// A box of cell (or face) indices [lo, hi], inclusive on both ends, with its location.
struct IndexBox {
    Vector3i vlo, vhi;
    VariableLocation varloc;

    IndexBox() = default;
    IndexBox(const Vector3i& lo, const Vector3i& hi, VariableLocation loc = VARLOC_CELL)
        : vlo(lo), vhi(hi), varloc(loc) {}

    const VariableLocation& type() const { return varloc; }
    const Vector3i& lo() const { return vlo; }
    const Vector3i& hi() const { return vhi; }
    int lo(int d) const { return vlo(d); }
    int hi(int d) const { return vhi(d); }
    int ilo() const { return vlo.x; }
    int jlo() const { return vlo.y; }
    int klo() const { return vlo.z; }
    int ihi() const { return vhi.x; }
    int jhi() const { return vhi.y; }
    int khi() const { return vhi.z; }
    int size(int d) const { return vhi(d) - vlo(d) + 1; }
    int stride() const { return size(0) * size(1) * size(2); }
    int capacity(int ncomp = 1) const { return stride() * ncomp; }
    int isValid() const { return vlo.x <= vhi.x && vlo.y <= vhi.y && vlo.z <= vhi.z; }
    int containsIndex(int i, int j, int k) const {
        return vlo.x <= i && i <= vhi.x && vlo.y <= j && j <= vhi.y && vlo.z <= k && k <= vhi.z;
    }
    // position of (i, j, k, comp) in an array laid out over this box, i fastest
    int offset(int i, int j, int k, int comp = 0) const {
        return (i - vlo.x) + size(0) * ((j - vlo.y) + size(1) * ((k - vlo.z) + size(2) * comp));
    }

    IndexBox& extend(int ilo_, int jlo_, int klo_, int ihi_, int jhi_, int khi_) {
        vlo.x -= ilo_;
        vlo.y -= jlo_;
        vlo.z -= klo_;
        vhi.x += ihi_;
        vhi.y += jhi_;
        vhi.z += khi_;
        return *this;
    }
    IndexBox& extend(int ie, int je, int ke) { return extend(ie, je, ke, ie, je, ke); }
    // every active direction of the driver's dimension
    IndexBox& extend(int e) { return extend(e, e, NDIM == 3 ? e : 0); }
    IndexBox& extendLo(int d, int n) { vlo(d) -= n; return *this; }
    IndexBox& extendHi(int d, int n) { vhi(d) += n; return *this; }
    IndexBox& shrinkLo(int d, int n) { return extendLo(d, -n); }
    IndexBox& shrinkHi(int d, int n) { return extendHi(d, -n); }
    // face box of direction d: one more index along d (no change when already staggered in d)
    IndexBox& staggerInDir(int d) {
        if (!varloc.isStaggered(d)) {
            varloc.setStaggered(d);
            vhi(d) += 1;
        }
        return *this;
    }
    int isCellBox() const { return varloc.isCellVar(); }
    int isFaceBox(int d) const { return varloc.isFaceVar(d); }
    int isFaceBox() const { return varloc.isFaceVar(); }

    static IndexBox Stagger(const IndexBox& cellbox, int d) { return IndexBox(cellbox).staggerInDir(d); }
    static IndexBox Extend(const IndexBox& b, int e) { return IndexBox(b).extend(e); }
};

inline bool operator==(const IndexBox& a, const IndexBox& b) {
    return a.vlo.x == b.vlo.x && a.vlo.y == b.vlo.y && a.vlo.z == b.vlo.z && a.vhi.x == b.vhi.x &&
           a.vhi.y == b.vhi.y && a.vhi.z == b.vhi.z && a.varloc.bits == b.varloc.bits;
}
inline bool operator!=(const IndexBox& a, const IndexBox& b) { return !(a == b); }

// Loop over every index of a box, k outermost (the reference's BEGIN/END_FOR_BOX_RANGE pair,
// SayakaBox.h:399-407).
#define BEGIN_FOR_BOX_RANGE(box, ii, jj, kk)                                             \
    {                                                                                  \
        const int ii##lo = (box).ilo(), ii##hi = (box).ihi();                           \
        const int jj##lo = (box).jlo(), jj##hi = (box).jhi();                           \
        const int kk##lo = (box).klo(), kk##hi = (box).khi();                           \
        for (int kk = kk##lo; kk <= kk##hi; ++kk) {                                     \
            for (int jj = jj##lo; jj <= jj##hi; ++jj) {                                 \
                for (int ii = ii##lo; ii <= ii##hi; ++ii) {
#define END_FOR_BOX_RANGE \
    }                     \
    }                     \
    }                     \
    }

// Unit increment along direction d, as a tuple (i, j, k) (SayakaCommons.h:40-55).
inline std::tuple<int, int, int> SelectStaggerIncr(int d) {
    return std::make_tuple(d == 0 ? 1 : 0, d == 1 ? 1 : 0, d == 2 ? 1 : 0);
}
inline void SelectStaggerIncr(int d, int& ii, int& jj, int& kk) {
    ii = d == 0;
    jj = d == 1;
    kk = d == 2;
}

// One of the 3^NDIM positions around a block, counted with i fastest; usable as the loop
// variable of `for (SurroundingIndex s = 0; s < SurroundingIndex::NumSurr; ++s)`.
struct SurroundingIndex {
    enum { NumSurr = NDIM == 3 ? 27 : 9, CenterIndex = NumSurr / 2 };
    int m_index = 0;
    SurroundingIndex(int index) : m_index(index) {}
    SurroundingIndex(int i, int j, int k) { setPos(i, j, k); }
    int ipos() const { return m_index % 3 - 1; }
    int jpos() const { return (m_index / 3) % 3 - 1; }
    int kpos() const { return NDIM == 3 ? m_index / 9 - 1 : 0; }
    void setPos(int i, int j, int k) { m_index = (i + 1) + 3 * (j + 1) + (NDIM == 3 ? 9 * (k + 1) : 0); }
    bool isCenter() const { return m_index == CenterIndex; }
    operator int() const { return m_index; }
    operator int&() { return m_index; }
};

// Face f = 2 * dir + side of a block; usable as a loop variable over FaceIndex::NumFace.
struct FaceIndex {
    enum { NumFace = 2 * NDIM };
    int m_index = 0;
    FaceIndex(int f) : m_index(f) {}
    FaceIndex(int dir, int side) : m_index(2 * dir + side) {}
    int dir() const { return m_index / 2; }
    int side() const { return m_index % 2; }
    operator int() const { return m_index; }
    operator int&() { return m_index; }
};

}  // namespace amd
}  // namespace sayaka
