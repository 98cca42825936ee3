// sayaka_types.h -- small value types of the host-side tree (dimension chosen at run time).
//
// The reference picks the dimension at compile time (SayakaCommons.h:4-37); this build keeps
// one library for 2D and 3D, so the dimension is a member of the tree and of every helper
// that needs it.  Index conventions are the reference's:
//   face  f = 2*dir + side                  (SayakaFaceIndex.h:90-118)
//   child c = iside + 2*jside + 4*kside     (SayakaChildIndex.h:199-216)
//   cell offset in a box: i fastest        (SayakaBox.h:214-217)
#pragma once

#include <array>
#include <cstdint>

namespace sayaka {
inline namespace amd {

constexpr int MAX_DIM = 3;
constexpr int MAX_FACE = 6;
constexpr int MAX_CHILD = 8;
constexpr int MAX_LEVEL = 16;
constexpr int PHYS_BNDRY = -20;  // neighbour id <= this is a physical boundary (SayakaTree.h:38-40)

enum TreeNodeType {
    NodeInvalid = 0, NodeIsLeaf = 1, NodeHasLeafChild = 2, NodeNoLeafChild = 3,
    // the reference's spellings (SayakaTree.h:27-32)
    TreeNodeType_Invalid = 0, TreeNodeType_IsLeaf = 1, TreeNodeType_HasLeafChild = 2, TreeNodeType_NoLeafChild = 3
};
enum TowerFaceType { FACE_FINE_FINE = 0, FACE_FINE_BC = 1, FACE_FINE_CRSE = 2, FACE_CRSE_FINE = 3 };
enum BCType { Neumann = 0, Dirichlet = 1, SimpleFill = 2 };

template <class T>
struct Vec3 {
    T x{}, y{}, z{};
    T& operator()(int d) { return d == 0 ? x : (d == 1 ? y : z); }
    const T& operator()(int d) const { return d == 0 ? x : (d == 1 ? y : z); }
    static Vec3 VecMake(T a, T b, T c) { return Vec3{a, b, c}; }
    void setZero() { x = y = z = T(0); }
};
using Vector3i = Vec3<int>;
using Vector3d = Vec3<double>;

inline int face_dir(int f) { return f / 2; }
inline int face_side(int f) { return f % 2; }
inline int face_opposite(int f) { return f ^ 1; }
inline int child_side(int c, int d) { return (c >> d) & 1; }

// Neighbour of child c across face f: the sibling index and whether it shares c's parent.
// Sibling across a face = flip the bit of the face direction; it is in-parent exactly when
// the face lies on the far side of c's position (SayakaChildIndex.cpp:300-331).
inline int child_neighbor(int c, int f, bool& same_parent) {
    const int d = face_dir(f);
    same_parent = child_side(c, d) != face_side(f);
    return c ^ (1 << d);
}

}  // namespace amd
}  // namespace sayaka
