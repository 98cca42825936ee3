// remap.h -- regrid data movement of the device fields.
#pragma once

#include "kernels.h"

namespace sydev {

// ------------------------------------------------------------------ regrid data movement
// After a regrid the device fields are still in the old numbering.  Block b of the new tree
// takes the values of block origin[b] of the old one (the reference packs its registered
// state data the same way, SayakaRegrid.cpp:530-591); blocks the regrid created (origin -1)
// are then filled level by level, coarsest first, from their parents by injection
// (initNewBlockDataAfterRegrid, SayakaRegrid.cpp:987-1105, with the default
// PROLONG_INJECTION, prolongCellBlockData_Injection, SayakaInterpPatch.cpp:155-187).
template <int D, int NB>
__global__ void __launch_bounds__((Geo<D, NB>::THREADS))
k_remap_gather(const int* __restrict__ origin, const double* __restrict__ src, double* __restrict__ dst) {
    using G = Geo<D, NB>;
    const int b = blockIdx.x;
    const int o = origin[b];
    if (o < 0) return;
    for (int c = threadIdx.x; c < G::NCELL; c += blockDim.x) dst[b * G::NCELL + c] = src[o * G::NCELL + c];
}

template <int D, int NB>
__global__ void __launch_bounds__((Geo<D, NB>::THREADS))
k_inject(const int* __restrict__ list, const BlockInfo* __restrict__ binfo, double* __restrict__ x) {
    using G = Geo<D, NB>;
    const int b = list[blockIdx.x];
    const int p = binfo[b].parent, wc = binfo[b].which_child;
    for (int c = threadIdx.x; c < G::NCELL; c += blockDim.x) {
        int i, j, k;
        G::cell_ijk(c, i, j, k);
        const int ic = (i >> 1) + side_of(wc, 0) * (NB / 2);
        const int jc = (j >> 1) + side_of(wc, 1) * (NB / 2);
        const int kc = D == 3 ? (k >> 1) + side_of(wc, 2) * (NB / 2) : 0;
        x[b * G::NCELL + c] = x[p * G::NCELL + G::cidx(ic, jc, kc)];
    }
}

}  // namespace sydev
