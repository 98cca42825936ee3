// statefill.h -- ghost fill of a user cell state with several ghost layers (the general fill
// the reference's drivers run on their state: FillPatch::fillBoundary, SayakaFillPatch.cpp:17-89).
//
// The state lives on the device as [comp][block][grown box], the block grown by ngrow ghost
// layers on every side (i fastest; 2D blocks are not grown in z).  One launch per tree level,
// coarsest first, one workgroup per block of the level (leaf or not), because a block's
// coarse/fine ghosts are interpolated from its parent's cells *and ghosts*.  Per block the
// grown box is staged in LDS and filled in the reference's three steps:
//   1. same-level surroundings (faces, edges, corners): copy of the neighbour's valid cells
//      (fillBlockBndryAtSameLevel, SayakaFillPatch.cpp:92-190);
//   2. coarse surroundings: prolongation from the parent's grown box -- injection,
//      centred-linear or MC-limited centred-linear per component
//      (fillBlockBndryFromCoarseLevel, SayakaFillPatch.cpp:192-299;
//      prolongCellBlockData_*, SayakaInterpPatch.cpp:157-391); every fine cell evaluates its
//      own coarse cell's slopes, which is the same arithmetic as the reference's per-coarse-
//      cell loop;
//   3. physical surroundings, in the order of fillBlockBndryCond (SayakaFillPatch.cpp:301-368):
//      reflection about the wall over all ngrow layers (BCPatch::fillBasicCellBlockBC,
//      SayakaBoundaryPatch.cpp:139-229), one LDS barrier per applied entry since an edge or
//      corner ghost can reflect a ghost filled just before.
// Steps 1-2 cover the nlayer ghost layers asked for; step 3 always covers all ngrow (the
// reference's BC boxes are built from the data's ngrow).  Only ghost cells are written back:
// other workgroups of the same level read this block's valid cells only.
#pragma once

#include "kernels.h"

namespace sydev {

constexpr int FILL_MAXCOMP = 16;
constexpr int FILL_NT = 256;

enum ProlongMethod { PR_INJECTION = 0, PR_CENTERED_LINEAR = 1, PR_CENTERED_LINEAR_LIMITED = 2 };

struct FillArgs {
    const int* list;        // blocks of the level
    const BlockInfo* binfo;
    const int* surr;        // [blk][27] SurroundingBlocks::blocks, [ii+1][jj+1][kk+1] order
    const int* bcref;       // [blk][27] SurroundingBlocks::bcrefs (SurroundingIndex, -1 none)
    const int* bctype;      // [nc][27][nphys] math BC of (surrounding, physical code), -1 none
    const double* bcval;    // [nc][27][nphys]
    const int* physcode;    // [nphys]
    double* state;          // component 0, block 0
    long long comp_stride;  // doubles per component
    int ngrow, nlayer, scomp, nc, nphys;
    int same_level_only;    // fillBlockBoundary(.., onlySameLevel = 1): no coarse surroundings
    int prolong[FILL_MAXCOMP];
};

__device__ inline int fill_surr_flat(int ii, int jj, int kk) { return (ii + 1) * 9 + (jj + 1) * 3 + (kk + 1); }

// IndexMapping::FineToCrse / CrseToFine (SayakaTreeData.h:34-44)
__device__ inline int fine_to_crse(int i, int off) { return (i + 64) / 2 - 32 + off; }

// limited_slope (SayakaInterpPatch.h:214-233): MC limiter, std::min / std::max semantics
__device__ inline double mc_slope(double ul, double uc, double ur) {
    const double gradc = 0.5 * (ur - ul);
    const double gradl = uc - ul;
    const double gradr = ur - uc;
    double signc = 0.0;
    if (gradc > 0) signc = 1.0;
    else if (gradc < 0) signc = -1.0;
    double grad = fabs(gradc);
    double t = 2.0 * signc * gradl;
    grad = t < grad ? t : grad;
    t = 2.0 * signc * gradr;
    grad = t < grad ? t : grad;
    grad = grad < 0.0 ? 0.0 : grad;
    grad *= signc;
    return grad;
}

template <int D, int NB>
__global__ void __launch_bounds__(FILL_NT) k_state_fill(FillArgs a) {
    extern __shared__ double tile[];
    __shared__ int s[27], r[27];
    const int b = a.list[blockIdx.x];
    const int g = a.ngrow, gz = D == 3 ? g : 0;
    constexpr int NZ = D == 3 ? NB : 1;
    const int GX = NB + 2 * g, GZ = NZ + 2 * gz;
    const int ntile = GX * GX * GZ;
    if (threadIdx.x < 27) {
        s[threadIdx.x] = a.surr[b * 27 + threadIdx.x];
        r[threadIdx.x] = a.bcref[b * 27 + threadIdx.x];
    }
    const int p = a.binfo[b].parent, wc = a.binfo[b].which_child;
    const int off[3] = {side_of(wc, 0) * (NB / 2), side_of(wc, 1) * (NB / 2), D == 3 ? side_of(wc, 2) * (NB / 2) : 0};
    const int nl = a.nlayer;
    auto tix = [&](int i, int j, int k) { return ((k + gz) * GX + (j + g)) * GX + (i + g); };
    __syncthreads();
    for (int cc = 0; cc < a.nc; cc++) {
        const int comp = a.scomp + cc;
        double* base = a.state + comp * a.comp_stride;
        double* blk = base + (long long)b * ntile;
        for (int e = threadIdx.x; e < ntile; e += FILL_NT) tile[e] = blk[e];
        __syncthreads();
        // steps 1 and 2: same-level copy and coarse prolongation within nlayer
        const int method = a.prolong[cc];
        for (int e = threadIdx.x; e < ntile; e += FILL_NT) {
            const int i = e % GX - g, j = (e / GX) % GX - g, k = e / (GX * GX) - gz;
            const int ii = i < 0 ? -1 : (i >= NB ? 1 : 0);
            const int jj = j < 0 ? -1 : (j >= NB ? 1 : 0);
            const int kk = k < 0 ? -1 : (k >= NZ ? 1 : 0);
            if (ii == 0 && jj == 0 && kk == 0) continue;
            if (i < -nl || i >= NB + nl || j < -nl || j >= NB + nl) continue;
            if (D == 3 && (k < -nl || k >= NZ + nl)) continue;
            const int sv = s[fill_surr_flat(ii, jj, kk)];
            if (sv >= 0) {
                tile[e] = base[(long long)sv * ntile + tix(i - ii * NB, j - jj * NB, k - kk * NZ)];
            } else if (sv > PHYS_BNDRY && p >= 0 && !a.same_level_only) {
                const double* par = base + (long long)p * ntile;
                const int ic = fine_to_crse(i, off[0]), jc = fine_to_crse(j, off[1]);
                const int kc = D == 3 ? fine_to_crse(k, off[2]) : k;
                const double c = par[tix(ic, jc, kc)];
                if (method == PR_INJECTION) {
                    tile[e] = c;
                } else {
                    double gr[3] = {0.0, 0.0, 0.0};
                    const int ci[3] = {ic, jc, kc};
#pragma unroll
                    for (int d = 0; d < D; d++) {
                        int up[3] = {ci[0], ci[1], ci[2]}, dn[3] = {ci[0], ci[1], ci[2]};
                        up[d]++;
                        dn[d]--;
                        const double vu = par[tix(up[0], up[1], up[2])];
                        const double vd = par[tix(dn[0], dn[1], dn[2])];
                        gr[d] = method == PR_CENTERED_LINEAR ? 0.5 * (vu - vd) : mc_slope(vd, c, vu);
                    }
                    // the fine cell's offset from its coarse cell centre, in coarse cells:
                    // -0.25 for the low half (CrseToFine(ic, 0, off) == i), +0.25 otherwise
                    const double irel = ((ic - off[0]) * 2 == i) ? -0.25 : 0.25;
                    const double jrel = ((jc - off[1]) * 2 == j) ? -0.25 : 0.25;
                    const double krel = D == 3 ? (((kc - off[2]) * 2 == k) ? -0.25 : 0.25) : 0.0;
                    tile[e] = c + irel * gr[0] + jrel * gr[1] + krel * gr[2];
                }
            }
        }
        __syncthreads();
        // step 3: physical boundary, sweep by sweep in the reference's order
        for (int sweep = 0; sweep < D; sweep++) {
            const int sweepdir = sweep % D;
            int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, inc[3] = {1, 1, 1};
            for (int d = 0; d <= sweep; d++) {
                lo[(d + sweep) % D] = -1;
                hi[(d + sweep) % D] = 1;
            }
            inc[sweepdir] = 2;
            for (int kk = lo[2]; kk <= hi[2]; kk += inc[2])
                for (int jj = lo[1]; jj <= hi[1]; jj += inc[1])
                    for (int ii = lo[0]; ii <= hi[0]; ii += inc[0]) {
                        const int q = fill_surr_flat(ii, jj, kk);
                        const int code = s[q];
                        if (code > PHYS_BNDRY) continue;   // uniform over the workgroup
                        int t = -1;
                        double bv = 0.0;
                        for (int h = 0; h < a.nphys; h++)
                            if (a.physcode[h] == code) {
                                t = a.bctype[(cc * 27 + q) * a.nphys + h];
                                bv = a.bcval[(cc * 27 + q) * a.nphys + h];
                            }
                        const int bnd[3] = {ii, jj, kk};
                        int rdir = sweepdir;
                        if (r[q] >= 0) {
                            const int rp[3] = {r[q] % 3 - 1, (r[q] / 3) % 3 - 1, D == 3 ? r[q] / 9 - 1 : 0};
                            rdir = -1;
                            for (int d = 0; d < 3; d++)
                                if (bnd[d] != rp[d]) rdir = d;
                        }
                        // a reference that is not a face of the surrounding would reflect out of
                        // the grown box (never the case for the tree's tables): no reflection
                        if (rdir >= 0 && bnd[rdir] == 0) rdir = -1;
                        const int n3[3] = {NB, NB, NZ};
                        int sg[3], of[3], blo[3], bhi[3];
                        const int ext[3] = {g, g, gz};
                        for (int d = 0; d < 3; d++) {
                            sg[d] = rdir == d ? -1 : 1;
                            of[d] = rdir != d ? 0 : (bnd[d] == -1 ? -1 : 2 * n3[d] - 1);
                            // the shifted valid box intersected with the ngrow layer box
                            blo[d] = bnd[d] == -1 ? -ext[d] : (bnd[d] == 1 ? n3[d] : 0);
                            bhi[d] = bnd[d] == -1 ? -1 : (bnd[d] == 1 ? n3[d] + ext[d] - 1 : n3[d] - 1);
                        }
                        const int nx = bhi[0] - blo[0] + 1, ny = bhi[1] - blo[1] + 1, nz = bhi[2] - blo[2] + 1;
                        if (t >= 0 && t <= 2)
                            for (int e = threadIdx.x; e < nx * ny * nz; e += FILL_NT) {
                                const int i = blo[0] + e % nx, j = blo[1] + (e / nx) % ny, k = blo[2] + e / (nx * ny);
                                const double vref = tile[tix(sg[0] * i + of[0], sg[1] * j + of[1], sg[2] * k + of[2])];
                                tile[tix(i, j, k)] = t == 0 ? vref : (t == 1 ? 2.0 * bv - vref : bv);
                            }
                        __syncthreads();
                    }
        }
        // ghost cells back to HBM
        for (int e = threadIdx.x; e < ntile; e += FILL_NT) {
            const int i = e % GX - g, j = (e / GX) % GX - g, k = e / (GX * GX) - gz;
            const bool inside = i >= 0 && i < NB && j >= 0 && j < NB && k >= 0 && k < NZ;
            if (!inside) blk[e] = tile[e];
        }
        __syncthreads();
    }
}

// ----------------------------------------------------------------------- face data
// The same fill on face data of direction fd (staggered in fd: one more face than cells along
// it), [comp][block][grown face box].  The reference fills face data only where no wall is near
// (fillBlockBndryCond exits for face data next to a wall, SayakaFillPatch.cpp:318-324): the host
// refuses trees with walls.
//   1. same-level surroundings over the staggered ranges (SayakaFillPatch.cpp:660-675).  These
//      overlap on the faces shared by two surroundings and include the block's own boundary
//      faces (fd index 0 and nb), which the reference overwrites with the neighbour's copy; the
//      surrounding later in the reference's loop order wins, i.e. the one with the larger fd
//      offset.  The reference fills the blocks of a level one after another, so where the two
//      copies of a shared face differ its result depends on its block order; for face data whose
//      shared faces agree (a synchronised face field) every order gives this result.
//   2. coarse surroundings one after another in the reference's order, over the ranges that
//      leave the block's own faces alone (dstNonOverlapRanges, SayakaFillPatch.cpp:677-700):
//      prolongFaceBlockData_Injection / _CenteredLinear (SayakaInterpPatch.cpp:447-591).  The
//      centred-linear form sets the fine faces on coarse faces first (tangential slopes), then
//      each face between two coarse faces to the mean of its fd neighbours, which can lie outside
//      the range; hence one surrounding at a time, with a barrier between the sweeps.
// Only ghost faces are written back.
struct FaceFillArgs {
    const int* list;
    const BlockInfo* binfo;
    const int* surr;        // [blk][27]
    double* state;          // component 0, block 0
    long long comp_stride;
    int fd, ngrow, nlayer, scomp, nc;
    int same_level_only;    // no coarse surroundings (fillBlockBoundary's onlySameLevel)
    int prolong[FILL_MAXCOMP];
};

template <int D, int NB>
__global__ void __launch_bounds__(FILL_NT) k_face_fill(FaceFillArgs a) {
    extern __shared__ double tile[];
    __shared__ int s[27];
    const int b = a.list[blockIdx.x];
    const int g = a.ngrow, gz = D == 3 ? g : 0, fd = a.fd;
    constexpr int NZ = D == 3 ? NB : 1;
    const int N3[3] = {NB, NB, NZ};
    const int V[3] = {NB + (fd == 0), NB + (fd == 1), NZ + (fd == 2)};   // valid face extents
    const int E0 = V[0] + 2 * g, E1 = V[1] + 2 * g, E2 = V[2] + 2 * gz;
    const int ntile = E0 * E1 * E2;
    if (threadIdx.x < 27) s[threadIdx.x] = a.surr[b * 27 + threadIdx.x];
    const int p = a.binfo[b].parent, wc = a.binfo[b].which_child;
    const int off[3] = {side_of(wc, 0) * (NB / 2), side_of(wc, 1) * (NB / 2), D == 3 ? side_of(wc, 2) * (NB / 2) : 0};
    const int nl = a.nlayer;
    auto tix = [&](int i, int j, int k) { return ((k + gz) * E1 + (j + g)) * E0 + (i + g); };
    __syncthreads();
    for (int cc = 0; cc < a.nc; cc++) {
        double* base = a.state + (a.scomp + cc) * a.comp_stride;
        double* blk = base + (long long)b * ntile;
        for (int e = threadIdx.x; e < ntile; e += FILL_NT) tile[e] = blk[e];
        __syncthreads();
        // 1. same-level copy
        for (int e = threadIdx.x; e < ntile; e += FILL_NT) {
            const int ijk[3] = {e % E0 - g, (e / E0) % E1 - g, e / (E0 * E1) - gz};
            int o[3], o2 = -2;   // o2: the fd offset of a second candidate (lower precedence)
            bool in_layer = true;
            for (int d = 0; d < 3; d++) {
                const int x = ijk[d], n = N3[d];
                if (d >= D) {
                    o[d] = 0;
                } else if (d != fd) {
                    o[d] = x < 0 ? -1 : (x >= n ? 1 : 0);
                    if (x < -nl || x > n - 1 + nl) in_layer = false;
                } else {
                    if (x < 0) o[d] = -1;
                    else if (x == 0) { o[d] = 0; o2 = -1; }
                    else if (x < n) o[d] = 0;
                    else if (x == n) { o[d] = 1; o2 = 0; }
                    else o[d] = 1;
                    if (x < -nl || x > n + nl) in_layer = false;
                }
            }
            if (!in_layer) continue;
            for (int t = 0; t < 2; t++) {
                int q[3] = {o[0], o[1], o[2]};
                if (t == 1) {
                    if (o2 == -2) break;
                    q[fd] = o2;
                }
                if (q[0] == 0 && q[1] == 0 && q[2] == 0) continue;
                const int sv = s[fill_surr_flat(q[0], q[1], q[2])];
                if (sv >= 0) {
                    tile[e] = base[(long long)sv * ntile + tix(ijk[0] - q[0] * NB, ijk[1] - q[1] * NB, ijk[2] - q[2] * NZ)];
                    break;
                }
            }
        }
        __syncthreads();
        // 2. coarse surroundings, one after another
        const int method = a.prolong[cc];
        if (p >= 0 && !a.same_level_only) {
            const double* par = base + (long long)p * ntile;
            for (int kk = (D == 3 ? -1 : 0); kk <= (D == 3 ? 1 : 0); kk++)
                for (int jj = -1; jj <= 1; jj++)
                    for (int ii = -1; ii <= 1; ii++) {
                        if (ii == 0 && jj == 0 && kk == 0) continue;
                        const int sv = s[fill_surr_flat(ii, jj, kk)];
                        if (!(sv < 0 && sv > PHYS_BNDRY)) continue;   // uniform over the workgroup
                        const int oo[3] = {ii, jj, kk};
                        int lo[3], hi[3];
                        for (int d = 0; d < 3; d++) {
                            const int n = N3[d];
                            if (d >= D) { lo[d] = 0; hi[d] = 0; continue; }
                            lo[d] = oo[d] < 0 ? -nl : (oo[d] > 0 ? n : 0);
                            hi[d] = oo[d] < 0 ? -1 : (oo[d] > 0 ? n + nl - 1 : n - 1);
                            if (d == fd) {
                                hi[d] += 1;                      // staggered
                                if (oo[d] == -1) hi[d] -= 1;     // non-overlapping
                                if (oo[d] == 1) lo[d] += 1;
                            }
                        }
                        const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
                        const int nbox = nx * ny * nz;
                        for (int e = threadIdx.x; e < nbox; e += FILL_NT) {
                            const int f[3] = {lo[0] + e % nx, lo[1] + (e / nx) % ny, lo[2] + e / (nx * ny)};
                            int c3[3];
                            for (int d = 0; d < 3; d++) c3[d] = d < D ? fine_to_crse(f[d], off[d]) : f[d];
                            const int side = (f[fd] + 64) & 1;
                            const double c = par[tix(c3[0], c3[1], c3[2])];
                            if (method == PR_INJECTION) {
                                double v = c;
                                if (side) {
                                    int u[3] = {c3[0], c3[1], c3[2]};
                                    u[fd]++;
                                    v = 0.5 * (c + par[tix(u[0], u[1], u[2])]);
                                }
                                tile[tix(f[0], f[1], f[2])] = v;
                            } else if (side == 0) {
                                double gr[3] = {0.0, 0.0, 0.0}, rel[3] = {0.0, 0.0, 0.0};
                                for (int d = 0; d < D; d++) {
                                    if (d == fd) continue;
                                    int up[3] = {c3[0], c3[1], c3[2]}, dn[3] = {c3[0], c3[1], c3[2]};
                                    up[d]++;
                                    dn[d]--;
                                    const double vu = par[tix(up[0], up[1], up[2])], vd = par[tix(dn[0], dn[1], dn[2])];
                                    gr[d] = method == PR_CENTERED_LINEAR ? 0.5 * (vu - vd) : mc_slope(vd, c, vu);
                                    rel[d] = ((c3[d] - off[d]) * 2 == f[d]) ? -0.25 : 0.25;
                                }
                                tile[tix(f[0], f[1], f[2])] = c + rel[0] * gr[0] + rel[1] * gr[1] + rel[2] * gr[2];
                            }
                        }
                        __syncthreads();
                        if (method != PR_INJECTION) {
                            for (int e = threadIdx.x; e < nbox; e += FILL_NT) {
                                const int f[3] = {lo[0] + e % nx, lo[1] + (e / nx) % ny, lo[2] + e / (nx * ny)};
                                if (((f[fd] + 64) & 1) == 0) continue;
                                int m[3] = {f[0], f[1], f[2]}, q[3] = {f[0], f[1], f[2]};
                                m[fd]--;
                                q[fd]++;
                                tile[tix(f[0], f[1], f[2])] = 0.5 * (tile[tix(m[0], m[1], m[2])] + tile[tix(q[0], q[1], q[2])]);
                            }
                            __syncthreads();
                        }
                    }
        }
        // ghost faces back to HBM
        for (int e = threadIdx.x; e < ntile; e += FILL_NT) {
            const int i = e % E0 - g, j = (e / E0) % E1 - g, k = e / (E0 * E1) - gz;
            const bool inside = i >= 0 && i < V[0] && j >= 0 && j < V[1] && k >= 0 && k < V[2];
            if (!inside) blk[e] = tile[e];
        }
        __syncthreads();
    }
}

// ----------------------------------------------------------------------- output packing
// Valid cells of a grown cell state in the order of a Chombo output file's data:datatype=0
// datasets (WriteDataHdf5, SayakaWriterHdf5.cpp:480-512): per block of the level (ascending
// id), per component, the block's valid cells, i fastest.  One launch over the blocks listed
// level by level; workgroup (q, c) = listed block q, component c.  A pure copy: 16 B of HBM
// traffic per cell.
template <int D, int NB>
__global__ void __launch_bounds__(256) k_pack_levels(const int* __restrict__ list, int ncomp,
                                                     const double* __restrict__ state, long long comp_stride,
                                                     int ngrow, double* __restrict__ out) {
    constexpr int NCELL = D == 3 ? NB * NB * NB : NB * NB;
    const int q = blockIdx.x / ncomp, c = blockIdx.x % ncomp;
    const int gx = NB + 2 * ngrow, gz = D == 3 ? ngrow : 0;
    const long long ntile = (long long)gx * gx * (D == 3 ? NB + 2 * ngrow : 1);
    const double* src = state + c * comp_stride + (long long)list[q] * ntile;
    double* dst = out + (long long)blockIdx.x * NCELL;
    for (int e = threadIdx.x; e < NCELL; e += 256) {
        const int i = e % NB, j = (e / NB) % NB, k = e / (NB * NB);
        dst[e] = src[((k + gz) * gx + (j + ngrow)) * gx + (i + ngrow)];
    }
}

// ----------------------------------------------------------------- registered-state remap
// A state registered with the tree through AmrTree::regrid and initNewBlockDataAfterRegrid
// (SayakaRegrid.cpp:530-591, 987-1105).  The host runs, on the new tree:
//   k_state_move      every block's whole grown box from its old position (origin), 0 for a
//                     new block (the packing moves whole blocks; new slots hold zeros);
//   then per level L holding new blocks, coarsest first:
//   k_state_fill / k_face_fill with same_level_only over the parents of level L's new blocks
//                     (fillBlockBoundary(.., onlySameLevel = 1): same level, physical BC);
//   k_prolong_new / k_prolong_new_face over level L's new blocks: the valid box from the
//                     parent (prolongBlock, SayakaFillPatch.cpp:464-483), and for face data
//                     the shared faces of existing same-level face neighbours
//                     (correctFaceBlockData, SayakaFillPatch.cpp:485-513).
__global__ void __launch_bounds__(256) k_state_move(const int* __restrict__ origin, const double* __restrict__ old,
                                                    long long old_cstride, double* __restrict__ st,
                                                    long long cstride, long long ntile) {
    const int b = blockIdx.x, c = blockIdx.y;
    const int o = origin[b];
    double* dst = st + c * cstride + (long long)b * ntile;
    if (o >= 0) {
        const double* src = old + c * old_cstride + (long long)o * ntile;
        for (long long e = threadIdx.x; e < ntile; e += 256) dst[e] = src[e];
    } else {
        for (long long e = threadIdx.x; e < ntile; e += 256) dst[e] = 0.0;
    }
}

struct ProlongNewArgs {
    const int* list;        // new blocks of one level
    const int* origin;      // per block: old id, -1 new
    const BlockInfo* binfo;
    double* state;
    long long comp_stride;
    int ngrow, nc, fd;
    int prolong[FILL_MAXCOMP];
};

// cell data: every valid cell of the new block from its parent's grown box, the arithmetic of
// the ghost prolongation in k_state_fill (prolongCellBlockData_*, SayakaInterpPatch.cpp:157-391)
template <int D, int NB>
__global__ void __launch_bounds__(FILL_NT) k_prolong_new(ProlongNewArgs a) {
    constexpr int NZ = D == 3 ? NB : 1;
    constexpr int NCELL = NB * NB * NZ;
    const int b = a.list[blockIdx.x];
    const int g = a.ngrow, gz = D == 3 ? g : 0;
    const int GX = NB + 2 * g;
    const long long ntile = (long long)GX * GX * (NZ + 2 * gz);
    const int p = a.binfo[b].parent, wc = a.binfo[b].which_child;
    const int off[3] = {side_of(wc, 0) * (NB / 2), side_of(wc, 1) * (NB / 2), D == 3 ? side_of(wc, 2) * (NB / 2) : 0};
    auto tix = [&](int i, int j, int k) { return ((k + gz) * GX + (j + g)) * GX + (i + g); };
    for (int cc = 0; cc < a.nc; cc++) {
        double* base = a.state + cc * a.comp_stride;
        const double* par = base + (long long)p * ntile;
        double* blk = base + (long long)b * ntile;
        const int method = a.prolong[cc];
        for (int e = threadIdx.x; e < NCELL; e += FILL_NT) {
            const int i = e % NB, j = (e / NB) % NB, k = e / (NB * NB);
            const int ic = fine_to_crse(i, off[0]), jc = fine_to_crse(j, off[1]);
            const int kc = D == 3 ? fine_to_crse(k, off[2]) : k;
            const double c = par[tix(ic, jc, kc)];
            double v = c;
            if (method != PR_INJECTION) {
                double gr[3] = {0.0, 0.0, 0.0};
                const int ci[3] = {ic, jc, kc};
#pragma unroll
                for (int d = 0; d < D; d++) {
                    int up[3] = {ci[0], ci[1], ci[2]}, dn[3] = {ci[0], ci[1], ci[2]};
                    up[d]++;
                    dn[d]--;
                    const double vu = par[tix(up[0], up[1], up[2])];
                    const double vd = par[tix(dn[0], dn[1], dn[2])];
                    gr[d] = method == PR_CENTERED_LINEAR ? 0.5 * (vu - vd) : mc_slope(vd, c, vu);
                }
                const double irel = ((ic - off[0]) * 2 == i) ? -0.25 : 0.25;
                const double jrel = ((jc - off[1]) * 2 == j) ? -0.25 : 0.25;
                const double krel = D == 3 ? (((kc - off[2]) * 2 == k) ? -0.25 : 0.25) : 0.0;
                v = c + irel * gr[0] + jrel * gr[1] + krel * gr[2];
            }
            blk[tix(i, j, k)] = v;
        }
    }
}

// face data of direction fd: the valid face box of the new block staged in LDS, prolonged
// (prolongFaceBlockData_*, SayakaInterpPatch.cpp:447-591; the centred-linear second sweep
// after a barrier), the faces shared with existing same-level face neighbours replaced by the
// neighbour's copy, then written back
template <int D, int NB>
__global__ void __launch_bounds__(FILL_NT) k_prolong_new_face(ProlongNewArgs a) {
    extern __shared__ double vt[];
    constexpr int NZ = D == 3 ? NB : 1;
    const int b = a.list[blockIdx.x];
    const int g = a.ngrow, gz = D == 3 ? g : 0, fd = a.fd;
    const int V[3] = {NB + (fd == 0), NB + (fd == 1), NZ + (fd == 2)};
    const int E0 = V[0] + 2 * g, E1 = V[1] + 2 * g, E2 = V[2] + 2 * gz;
    const long long ntile = (long long)E0 * E1 * E2;
    const int nv = V[0] * V[1] * V[2];
    const int p = a.binfo[b].parent, wc = a.binfo[b].which_child;
    const int off[3] = {side_of(wc, 0) * (NB / 2), side_of(wc, 1) * (NB / 2), D == 3 ? side_of(wc, 2) * (NB / 2) : 0};
    auto tix = [&](int i, int j, int k) { return ((k + gz) * E1 + (j + g)) * E0 + (i + g); };
    auto vix = [&](int i, int j, int k) { return (k * V[1] + j) * V[0] + i; };
    for (int cc = 0; cc < a.nc; cc++) {
        double* base = a.state + cc * a.comp_stride;
        const double* par = base + (long long)p * ntile;
        double* blk = base + (long long)b * ntile;
        const int method = a.prolong[cc];
        for (int e = threadIdx.x; e < nv; e += FILL_NT) {
            const int f[3] = {e % V[0], (e / V[0]) % V[1], e / (V[0] * V[1])};
            int c3[3];
            for (int d = 0; d < 3; d++) c3[d] = d < D ? fine_to_crse(f[d], off[d]) : f[d];
            const int side = (f[fd] + 64) & 1;
            const double c = par[tix(c3[0], c3[1], c3[2])];
            if (method == PR_INJECTION) {
                double v = c;
                if (side) {
                    int u[3] = {c3[0], c3[1], c3[2]};
                    u[fd]++;
                    v = 0.5 * (c + par[tix(u[0], u[1], u[2])]);
                }
                vt[e] = v;
            } else if (side == 0) {
                double gr[3] = {0.0, 0.0, 0.0}, rel[3] = {0.0, 0.0, 0.0};
                for (int d = 0; d < D; d++) {
                    if (d == fd) continue;
                    int up[3] = {c3[0], c3[1], c3[2]}, dn[3] = {c3[0], c3[1], c3[2]};
                    up[d]++;
                    dn[d]--;
                    const double vu = par[tix(up[0], up[1], up[2])], vd = par[tix(dn[0], dn[1], dn[2])];
                    gr[d] = method == PR_CENTERED_LINEAR ? 0.5 * (vu - vd) : mc_slope(vd, c, vu);
                    rel[d] = ((c3[d] - off[d]) * 2 == f[d]) ? -0.25 : 0.25;
                }
                vt[e] = c + rel[0] * gr[0] + rel[1] * gr[1] + rel[2] * gr[2];
            }
        }
        __syncthreads();
        if (method != PR_INJECTION) {
            for (int e = threadIdx.x; e < nv; e += FILL_NT) {
                const int f[3] = {e % V[0], (e / V[0]) % V[1], e / (V[0] * V[1])};
                if (((f[fd] + 64) & 1) == 0) continue;
                int m[3] = {f[0], f[1], f[2]}, q[3] = {f[0], f[1], f[2]};
                m[fd]--;
                q[fd]++;
                vt[e] = 0.5 * (vt[vix(m[0], m[1], m[2])] + vt[vix(q[0], q[1], q[2])]);
            }
            __syncthreads();
        }
        for (int side = 0; side < 2; side++) {
            const int j = a.binfo[b].nbr[2 * fd + side];
            if (j < 0 || a.origin[j] < 0) continue;   // uniform over the workgroup
            const double* nb_ = base + (long long)j * ntile;
            const int dl = side == 0 ? 0 : V[fd] - 1, sl = side == 0 ? V[fd] - 1 : 0;
            const int t0 = fd == 0 ? 1 : 0, t1 = fd == 2 ? 1 : 2;   // the two tangential directions
            const int n0 = V[t0], n1 = V[t1];
            for (int e = threadIdx.x; e < n0 * n1; e += FILL_NT) {
                int dsti[3], srci[3];
                dsti[fd] = dl;
                srci[fd] = sl;
                dsti[t0] = srci[t0] = e % n0;
                dsti[t1] = srci[t1] = e / n0;
                vt[vix(dsti[0], dsti[1], dsti[2])] = nb_[tix(srci[0], srci[1], srci[2])];
            }
            __syncthreads();
        }
        for (int e = threadIdx.x; e < nv; e += FILL_NT) {
            const int f[3] = {e % V[0], (e / V[0]) % V[1], e / (V[0] * V[1])};
            blk[tix(f[0], f[1], f[2])] = vt[e];
        }
        __syncthreads();
    }
}

}  // namespace sydev
