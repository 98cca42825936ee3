// dev_common.h -- device-side data model of the block-structured MAC solve (CDNA4, gfx950).
//
// Every field lives in HBM as one packed array of fixed-size blocks, indexed by the host
// tree's block id: cell fields [block][k][j][i] (nb^d values, i fastest), face fields
// [block][dir][k][j][i] over the cell box staggered in dir.  No ghost cells are stored:
// a kernel stages its block plus one ghost layer in LDS, gathering the ghosts from the
// neighbour blocks through the per-block index table below (same-level neighbour, or the
// coarse neighbour for coarse/fine faces), which is exactly the value the reference's
// two-stage ghost fill (same-level copy, then injection from the parent's ghost) produces.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sydev {

enum FaceType { FT_FF = 0, FT_BC = 1, FT_FC = 2, FT_CF = 3 };  // SayakaTower.h:15-20
enum BCKind { BC_NEUMANN = 0, BC_DIRICHLET = 1, BC_SIMPLEFILL = 2 };

constexpr int PHYS_BNDRY = -20;

// Per-block index table, built by the host from the tree after every regrid.
struct BlockInfo {
    int level;
    int leaf;         // 1 for leaf blocks
    int parent;
    int which_child;  // child index within the parent
    int nbr[6];       // same-level face neighbour; -1 coarse side; <= -20 physical wall
    int cnbr[6];      // for nbr == -1: the coarse leaf across the face (parent's neighbour)
    int nbr_leaf[6];  // leaf flag of nbr[f] when nbr[f] >= 0
    int bct[6];       // math BC on wall faces (BCKind), -1 elsewhere
    int child[8];
    int nfc[6][4];    // children of the same-level neighbour across face f that touch f,
                      // indexed by their tangential sides s1 + 2*s2 (-1 when it has none)
};

struct BlockMetric {
    double dx[3];
    double ds[3];  // dv / dx[d]  (SayakaTree.h:293-301)
    double dv;
    double pad;
};

template <int D, int NB>
struct Geo {
    static constexpr int NBZ = D == 3 ? NB : 1;
    static constexpr int NCELL = NB * NB * NBZ;
    static constexpr int NF = 2 * D;
    static constexpr int NFACE = (NB + 1) * NB * NBZ / (D == 3 ? 1 : 1);  // per direction
    static constexpr int NSLAB = NB * NBZ;                                // cells on one face
    static constexpr int GX = NB + 2, GY = NB + 2, GZ = D == 3 ? NB + 2 : 1;
    static constexpr int KG = D == 3 ? 1 : 0;
    static constexpr int NTILE = GX * GY * GZ;
    static constexpr int NGHOST = NF * NSLAB;
    static constexpr int THREADS = NCELL < 512 ? (NCELL < 64 ? 64 : NCELL) : 512;
    static constexpr int NCHILD = 1 << D;

    __host__ __device__ static constexpr int tidx(int i, int j, int k) {
        return (i + 1) + GX * ((j + 1) + GY * (k + KG));
    }
    __host__ __device__ static constexpr int cidx(int i, int j, int k) { return i + NB * (j + NB * k); }
    // extents of the face box of direction d
    __host__ __device__ static constexpr int fx(int d) { return NB + (d == 0); }
    __host__ __device__ static constexpr int fy(int d) { return NB + (d == 1); }
    __host__ __device__ static constexpr int fidx(int d, int i, int j, int k) {
        return i + fx(d) * (j + fy(d) * k);
    }
    __host__ __device__ static void cell_ijk(int c, int& i, int& j, int& k) {
        i = c % NB;
        j = (c / NB) % NB;
        k = c / (NB * NB);
    }
};

// Tangential coordinates of a face slab: for face direction d the two other axes in
// increasing order (a1 < a2); in 2D a2 is the inactive z axis.
template <int D>
__host__ __device__ inline void tang_axes(int d, int& a1, int& a2) {
    if (D == 3) {
        a1 = d == 0 ? 1 : 0;
        a2 = d == 2 ? 1 : 2;
    } else {
        a1 = d == 0 ? 1 : 0;
        a2 = 2;
    }
}

__device__ inline int side_of(int child, int d) { return (child >> d) & 1; }

}  // namespace sydev
