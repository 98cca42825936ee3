// coef.h -- device set-up of the operator coefficients (once per solve / regrid).
//
// The reference keeps B and the IB fractions per block (face arrays with both boundary
// faces, a cell array for vf), synchronises the coarse side of every coarse/fine face to
// the average of the fine faces, and averages everything down to the parent blocks
// (SayakaMacCoef.cpp:319-367, SayakaMacIB.cpp:69-96, SayakaTower.cpp:298-415).  The face
// weights w = beta*ds/(dx*kappa) * B * a of cell_eval_coef (SayakaMacOperator.cpp:1115-1167,
// 1256-1266) are then precomputed per block face, so the stencil kernels read one weight
// per face instead of B, a and the metric.
#pragma once

#include "dev_common.h"

namespace sydev {

// Fine blocks' FINE_CRSE faces -> parent's boundary face sub-box, average of the
// 2^(d-1) fine faces (restrictFaceBlockData_Average, SayakaInterpPatch.cpp:627-691).
template <int D, int NB>
__global__ void k_sync_fine(const int* list, const BlockInfo* binfo, double* f0, double* f1) {
    using G = Geo<D, NB>;
    const int b = list[blockIdx.x];
    const int d = blockIdx.y % D;                          // direction; array f0 or f1
    double* fdata = (int)blockIdx.y < D ? f0 : f1;
    __shared__ BlockInfo bi;
    if (threadIdx.x == 0) bi = binfo[b];
    __syncthreads();
    int a1, a2;
    tang_axes<D>(d, a1, a2);
    const int h = NB / 2;
    const int nq = D == 3 ? h * h : h;
    const double scale = 1.0 / ((2 * 2 * (D == 3 ? 2 : 1)) / 2);
    for (int s = 0; s <= 1; s++) {
        if (!(bi.nbr[2 * d + s] < 0 && bi.nbr[2 * d + s] > PHYS_BNDRY)) continue;
        const int p = bi.parent, wc = bi.which_child;
        const int ext = d == 2 ? G::NBZ : NB;
        for (int q = threadIdx.x; q < nq; q += blockDim.x) {
            const int u1 = q % h, u2 = D == 3 ? q / h : 0;
            int cq[3] = {0, 0, 0};
            cq[d] = s ? ext : 0;
            cq[a1] = u1 + side_of(wc, a1) * h;
            if (D == 3) cq[a2] = u2 + side_of(wc, a2) * h;
            double val = 0.0;
            for (int q2 = 0; q2 <= (D == 3 ? 1 : 0); q2++)
                for (int q1 = 0; q1 <= 1; q1++) {
                    int fq[3] = {0, 0, 0};
                    fq[d] = s ? ext : 0;
                    fq[a1] = 2 * u1 + q1;
                    if (D == 3) fq[a2] = 2 * u2 + q2;
                    val += fdata[(b * D + d) * G::NFACE + G::fidx(d, fq[0], fq[1], fq[2])];
                }
            fdata[(p * D + d) * G::NFACE + G::fidx(d, cq[0], cq[1], cq[2])] = val * scale;
        }
    }
}

// Coarse leaves' CRSE_FINE faces <- the neighbour's opposite face slab (SayakaTower.cpp:388-413).
template <int D, int NB>
__global__ void k_sync_crse(const int* list, const BlockInfo* binfo, double* f0, double* f1, int m) {
    using G = Geo<D, NB>;
    const int b = list[blockIdx.x];
    const int d = blockIdx.y % D;
    double* fdata = (int)blockIdx.y < D ? f0 : f1;
    __shared__ BlockInfo bi;
    if (threadIdx.x == 0) bi = binfo[b];
    __syncthreads();
    int a1, a2;
    tang_axes<D>(d, a1, a2);
    const int ext = d == 2 ? G::NBZ : NB;
    for (int s = 0; s <= 1; s++) {
        const int f = 2 * d + s;
        if (!(bi.nbr[f] >= 0 && bi.level < m && !bi.nbr_leaf[f])) continue;
        const int n = bi.nbr[f];
        for (int q = threadIdx.x; q < G::NSLAB; q += blockDim.x) {
            int o[3] = {0, 0, 0}, r[3] = {0, 0, 0};
            o[a1] = r[a1] = q % NB;
            if (D == 3) o[a2] = r[a2] = q / NB;
            o[d] = s ? ext : 0;
            r[d] = s ? 0 : ext;
            fdata[(b * D + d) * G::NFACE + G::fidx(d, o[0], o[1], o[2])] =
                fdata[(n * D + d) * G::NFACE + G::fidx(d, r[0], r[1], r[2])];
        }
    }
}

// Face data of non-leaf blocks of one level <- average of the children's faces
// (FillPatch::restrictBlock, SayakaFillPatch.cpp:420-447).  Children are processed in index
// order there, so on the middle plane the child on the high side wins.
template <int D, int NB>
__global__ void k_restrict_face(const int* list, const BlockInfo* binfo, double* f0, double* f1) {
    using G = Geo<D, NB>;
    const int b = list[blockIdx.x];
    const int d = blockIdx.y % D;
    double* fdata = (int)blockIdx.y < D ? f0 : f1;
    const int h = NB / 2;
    const double scale = 1.0 / ((2 * 2 * (D == 3 ? 2 : 1)) / 2);
    int a1, a2;
    tang_axes<D>(d, a1, a2);
    for (int q = threadIdx.x; q < G::NFACE; q += blockDim.x) {
        int c[3];
        c[0] = q % G::fx(d);
        c[1] = (q / G::fx(d)) % G::fy(d);
        c[2] = q / (G::fx(d) * G::fy(d));
        int side[3] = {0, 0, 0}, off[3] = {0, 0, 0};
        for (int a = 0; a < D; a++) {
            side[a] = c[a] >= h ? 1 : 0;
            off[a] = side[a] * h;
        }
        const int ch = side[0] | (side[1] << 1) | (side[2] << 2);
        const int fb = binfo[b].child[ch];
        double val = 0.0;
        for (int q2 = 0; q2 <= (D == 3 ? 1 : 0); q2++)
            for (int q1 = 0; q1 <= 1; q1++) {
                int fq[3] = {0, 0, 0};
                fq[d] = (c[d] - off[d]) * 2;
                fq[a1] = (c[a1] - off[a1]) * 2 + q1;
                if (D == 3) fq[a2] = (c[a2] - off[a2]) * 2 + q2;
                val += fdata[(fb * D + d) * G::NFACE + G::fidx(d, fq[0], fq[1], fq[2])];
            }
        fdata[(b * D + d) * G::NFACE + q] = val * scale;
    }
}

// Cell data of non-leaf blocks <- plain average of the 2^d children cells
// (restrictCellBlockData_Average, SayakaInterpPatch.cpp:398-444).
template <int D, int NB>
__global__ void k_restrict_cell_avg(const int* list, const BlockInfo* binfo, double* x) {
    using G = Geo<D, NB>;
    const int b = list[blockIdx.x];
    const double wgt = 1.0 / G::NCHILD;
    const int h = NB / 2;
    for (int c = threadIdx.x; c < G::NCELL; c += blockDim.x) {
        int i, j, k;
        G::cell_ijk(c, i, j, k);
        const int si = i >= h, sj = j >= h, sk = (D == 3 && k >= h);
        const int fb = binfo[b].child[si | (sj << 1) | (sk << 2)];
        const int fi0 = (i - si * h) * 2, fj0 = (j - sj * h) * 2, fk0 = D == 3 ? (k - sk * h) * 2 : 0;
        double sum = 0.0;
        for (int kk = 0; kk <= (D == 3 ? 1 : 0); kk++)
            for (int jj = 0; jj <= 1; jj++)
                for (int ii = 0; ii <= 1; ii++) sum += x[fb * G::NCELL + G::cidx(fi0 + ii, fj0 + jj, fk0 + kk)];
        x[b * G::NCELL + c] = sum * wgt;
    }
}

// Face weights of every block for its own tower level, plus the coarse-side weight
// (kappa = 0.75) of each boundary face whose same-level neighbour exists, used where that
// face is CRSE_FINE in a finer tower (SayakaMacOperator.cpp:1139-1165).
template <int D, int NB>
__global__ void k_weights(const int* list, const BlockInfo* binfo, const BlockMetric* met, const double* bco,
                          const double* af, double beta, double* w, double* wcf) {
    using G = Geo<D, NB>;
    const int b = list[blockIdx.x];
    __shared__ BlockInfo bi;
    if (threadIdx.x == 0) bi = binfo[b];
    __syncthreads();
    const BlockMetric mt = met[b];
    for (int d = 0; d < D; d++) {
        const int ext = d == 2 ? G::NBZ : NB;
        int a1, a2;
        tang_axes<D>(d, a1, a2);
        for (int q = threadIdx.x; q < G::NFACE; q += blockDim.x) {
            int c[3];
            c[0] = q % G::fx(d);
            c[1] = (q / G::fx(d)) % G::fy(d);
            c[2] = q / (G::fx(d) * G::fy(d));
            const int fi = (b * D + d) * G::NFACE + q;
            const double B = bco[fi], A = af[fi];
            const double dsd = mt.ds[d], dxd = mt.dx[d];
            double off;
            const int n = c[d];
            if (n == 0 || n == ext) {
                const int f = 2 * d + (n == ext ? 1 : 0);
                const int nb = bi.nbr[f];
                if (nb <= PHYS_BNDRY) off = bi.bct[f] == BC_NEUMANN ? 0.0 : beta * dsd / dxd;
                else if (nb < 0) off = beta * dsd / (dxd * 1.5);
                else {
                    off = beta * dsd / dxd;
                    double oc = beta * dsd / (dxd * 0.75);
                    oc *= B;
                    oc *= A;
                    const int slab = c[a1] + NB * (D == 3 ? c[a2] : 0);
                    wcf[(b * G::NF + f) * G::NSLAB + slab] = oc;
                }
            } else {
                off = beta * dsd / dxd;
            }
            off *= B;
            off *= A;
            w[fi] = off;
        }
    }
}

// fixBlockBC (SayakaMacOperator.cpp:1501-1594): rhs += 2*coef*phi_bc / (dv*vf) on wall cells.
template <int D, int NB>
__global__ void k_fix_bc(const int* list, const BlockInfo* binfo, const BlockMetric* met, const double* bco,
                         const double* af, const double* vf, double beta, const double* phi, int ng,
                         double* rhs) {
    using G = Geo<D, NB>;
    const int b = list[blockIdx.x];
    __shared__ BlockInfo bi;
    if (threadIdx.x == 0) bi = binfo[b];
    __syncthreads();
    const BlockMetric mt = met[b];
    const int lx = NB + 2 * ng, ly = NB + 2 * ng, kg = D == 3 ? ng : 0, lz = G::NBZ + 2 * kg;
    const int gbase = b * lx * ly * lz;
    auto P = [&](int i, int j, int k) { return phi[gbase + (i + ng) + lx * ((j + ng) + ly * (k + kg))]; };
    for (int f = 0; f < G::NF; f++) {
        if (bi.nbr[f] > PHYS_BNDRY) continue;
        const int bct = bi.bct[f];
        const int d = f >> 1, s = f & 1;
        int a1, a2;
        tang_axes<D>(d, a1, a2);
        const int ext = d == 2 ? G::NBZ : NB;
        for (int q = threadIdx.x; q < G::NSLAB; q += blockDim.x) {
            int c[3] = {0, 0, 0};
            c[a1] = q % NB;
            if (D == 3) c[a2] = q / NB;
            c[d] = s ? ext - 1 : 0;
            int g[3] = {c[0], c[1], c[2]};
            g[d] += s ? 1 : -1;
            int fq[3] = {c[0], c[1], c[2]};
            fq[d] += s;
            const double phi0 = P(c[0], c[1], c[2]);
            const double phi1 = P(g[0], g[1], g[2]);
            const int fi = (b * D + d) * G::NFACE + G::fidx(d, fq[0], fq[1], fq[2]);
            double coef = beta * mt.ds[d] / mt.dx[d];
            coef *= bco[fi];
            coef *= af[fi];
            double corr = 0.0;
            if (bct == BC_DIRICHLET) {
                const double phibc = 0.5 * (phi0 + phi1);
                corr = phibc * 2.0 * coef;
            } else if (bct == BC_SIMPLEFILL) {
                corr = phi1 * 2.0 * coef;
            }
            const int ci = b * G::NCELL + G::cidx(c[0], c[1], c[2]);
            const double v = vf[ci];
            if (v > 0) {
                corr /= (mt.dv * v);
                rhs[ci] += corr;
            }
        }
        __syncthreads();  // corner cells are touched by two faces: keep the face order
    }
}

}  // namespace sydev
