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
            } else if (sv > PHYS_BNDRY && p >= 0) {
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

}  // namespace sydev
