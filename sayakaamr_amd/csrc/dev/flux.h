// flux.h -- face flux of the solution (the projection's pressure gradient).
//
// MacSolver::mg_flux_level with the ghost-fill path (SayakaMacOperator.cpp:161-203) and
// mg_flux_block_usefill (929-1025): on the faces of direction d of every block of tower m,
//   flux = -(phi(i) - phi(i-1)) / dh                      in-block and same-level faces,
//   0                                                     wall faces,
//   -(phi(i) - phi(i-1)) / (1.5 dh), the coarse ghost     fine side of a coarse/fine face,
//   the average of the 2^(d-1) fine-side fluxes           coarse side (syncSubLevelCrseFineFlux
//                                                         with FLUX_SYNC_AVG, SayakaTower.cpp:
//                                                         298-415, restrictFaceBlockData_Average,
//                                                         SayakaInterpPatch.cpp:627-691).
// The reference computes the sub-levels from fine to coarse so that the fine fluxes exist when
// the coarse side copies their average.  Here the coarse side recomputes each fine flux from
// the two cells it joins (the fine cell and this block's own cell, which is what the fine
// block's injected ghost holds), with the fine block's own dh and in the reference's summation
// order, so every block of the tower is independent: one launch, bit-identical values.  The
// reference also leaves the restricted fine fluxes in the faces of the non-leaf parents; those
// blocks are outside the tower and are not written here.
#pragma once

#include "kernels.h"

namespace sydev {

struct FluxArgs {
    const int* list;          // tower m
    int nlist;
    int m;
    const BlockInfo* binfo;
    const BlockMetric* met;
    const double* x;          // phi, [blk][cell]
    int dir;
    double* out;              // [blk][face of dir]
};

template <int D, int NB>
__global__ void __launch_bounds__((Geo<D, NB>::THREADS)) k_flux(FluxArgs a) {
    using G = Geo<D, NB>;
    __shared__ int biw[BI_WORDS];
    const int b = a.list[blockIdx.x];
    if ((int)threadIdx.x < BI_WORDS) biw[threadIdx.x] = reinterpret_cast<const int*>(a.binfo + b)[threadIdx.x];
    __syncthreads();
    const BlockInfo& bi = *reinterpret_cast<const BlockInfo*>(biw);
    const int d = a.dir;
    const double dh = a.met[b].dx[d];
    const double dh_fine_crse = dh * 1.5;
    const int fx = G::fx(d), fy = G::fy(d);
    const int ext = d == 2 ? G::NBZ : NB;
    int a1, a2;
    tang_axes<D>(d, a1, a2);
    const double* xb = a.x + b * G::NCELL;
    for (int fi = threadIdx.x; fi < G::NFACE; fi += blockDim.x) {
        int q[3] = {fi % fx, (fi / fx) % fy, fi / (fx * fy)};
        const int ijk = q[d];
        double v;
        if (ijk > 0 && ijk < ext) {
            int ql[3] = {q[0], q[1], q[2]};
            ql[d] -= 1;
            v = -(xb[G::cidx(q[0], q[1], q[2])] - xb[G::cidx(ql[0], ql[1], ql[2])]) / dh;
        } else {
            const int s = ijk == ext ? 1 : 0;
            const int f = 2 * d + s;
            const int ft = face_type<D, NB>(bi, f, a.m);
            int qo[3] = {q[0], q[1], q[2]};
            qo[d] = s ? ext - 1 : 0;
            const double own = xb[G::cidx(qo[0], qo[1], qo[2])];
            if (ft == FT_BC) {
                v = 0.0;
            } else if (ft == FT_FF) {
                int r[3] = {q[0], q[1], q[2]};
                r[d] = s ? 0 : ext - 1;
                const double g = a.x[bi.nbr[f] * G::NCELL + G::cidx(r[0], r[1], r[2])];
                v = -(s ? g - own : own - g) / dh;
            } else if (ft == FT_FC) {
                // the coarse leaf cell the injected ghost holds (SayakaFillPatch.cpp:192-299)
                const int wc = bi.which_child;
                int r[3] = {0, 0, 0};
                r[d] = s ? 0 : ext - 1;
                r[a1] = (q[a1] >> 1) + side_of(wc, a1) * (NB / 2);
                if (D == 3) r[a2] = (q[a2] >> 1) + side_of(wc, a2) * (NB / 2);
                const double g = a.x[bi.cnbr[f] * G::NCELL + G::cidx(r[0], r[1], r[2])];
                v = -(s ? g - own : own - g) / dh_fine_crse;
            } else {
                // coarse side: average of the fine fluxes across this face cell, each from the
                // fine cell and this cell over the fine block's 1.5 dh (kind, jind, iind order)
                const int half = NB / 2;
                const int t1 = q[a1], t2 = D == 3 ? q[a2] : 0;
                const int s1 = t1 >= half ? 1 : 0, s2 = (D == 3 && t2 >= half) ? 1 : 0;
                const int fb = bi.nfc[f][s1 + 2 * s2];
                const double dhf = a.met[fb].dx[d] * 1.5;
                double val = 0;
                for (int q2 = 0; q2 <= (D == 3 ? 1 : 0); q2++)
                    for (int q1 = 0; q1 <= 1; q1++) {
                        int c[3] = {0, 0, 0};
                        c[d] = s ? 0 : ext - 1;
                        c[a1] = (t1 - s1 * half) * 2 + q1;
                        if (D == 3) c[a2] = (t2 - s2 * half) * 2 + q2;
                        const double fine = a.x[fb * G::NCELL + G::cidx(c[0], c[1], c[2])];
                        // the fine block's face: its low face when it lies on this block's high
                        // side (phi(0) - ghost), its high face otherwise (ghost - phi(ext-1))
                        val += -(s ? fine - own : own - fine) / dhf;
                    }
                v = val * (D == 3 ? 0.25 : 0.5);
            }
        }
        a.out[b * G::NFACE + fi] = v;
    }
}

// The one-layer face ghosts that mg_flux_level leaves in its phi argument: mg_fillbndry_level
// (SayakaMacUtil.cpp:210-251) fills every block of the visited levels with injection and
// neither corners nor physical boundaries (FillPatch::fillLevelBoundary, SayakaFillPatch.cpp:
// 34-299): a same-level neighbour gives its boundary cell (leaf or not), a coarse neighbour the
// cell of the parent's neighbour that the injection reads through the parent's ghost.  One
// workgroup per block of the list, out[block][face][slab] with the lower tangential axis
// fastest; wall faces are written as 0 (the fill leaves those ghosts alone).
template <int D, int NB>
__global__ void __launch_bounds__(256) k_solver_ghosts(const int* list, const BlockInfo* binfo, const double* x,
                                                     double* out) {
    using G = Geo<D, NB>;
    const int b = list[blockIdx.x];
    const BlockInfo& bi = binfo[b];
    for (int g = threadIdx.x; g < G::NGHOST; g += 256) {
        const int f = g / G::NSLAB, p = g % G::NSLAB;
        const int d = f >> 1, s = f & 1;
        int a1, a2;
        tang_axes<D>(d, a1, a2);
        const int ext = d == 2 ? G::NBZ : NB;
        int q[3] = {0, 0, 0};
        q[a1] = p % NB;
        if (D == 3) q[a2] = p / NB;
        const int n = bi.nbr[f];
        double v = 0.0;
        if (n >= 0) {
            q[d] = s ? 0 : ext - 1;
            v = x[n * G::NCELL + G::cidx(q[0], q[1], q[2])];
        } else if (n > PHYS_BNDRY) {
            int r[3] = {0, 0, 0};
            r[d] = s ? 0 : ext - 1;
            r[a1] = (q[a1] >> 1) + side_of(bi.which_child, a1) * (NB / 2);
            if (D == 3) r[a2] = (q[a2] >> 1) + side_of(bi.which_child, a2) * (NB / 2);
            v = x[bi.cnbr[f] * G::NCELL + G::cidx(r[0], r[1], r[2])];
        }
        out[(size_t)b * G::NGHOST + g] = v;
    }
}

}  // namespace sydev
