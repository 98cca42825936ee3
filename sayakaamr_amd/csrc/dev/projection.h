// projection.h -- the projection step of a flow solver, on the device.
//
// The reference's flow demo (Test7.h:580-657) builds the MAC rhs from the divergence of the
// face velocity (calc_pres_rhs: rhs = -div u / dt), solves for the pressure and corrects the
// velocity with mg_flux_level's flux (corr_vel: u += dt flux), then synchronises coarse/fine
// face velocities (fill_vel_all).  Here rhs = +div u and u -= B flux, i.e. the demo with dt = 1
// and the pressure's sign flipped, with the variable coefficient B of the solve.  These
// kernels work on the packed device layout:
//   k_face_velocity  a seeded U(-1,1) face velocity keyed on (seed, level, global face index,
//                    dir) -- the definition of sayakaamr_amd/workloads.py:face_velocity, so a
//                    face shared by two blocks of one level gets one value;
//   k_divergence     div = sum_d (a u ds)_hi - (a u ds)_lo, over dv vf on fluid cells (the
//                    definition of workloads.py:divergence_rhs), with per-block sums of
//                    dv vf div and dv vf over the leaf blocks for the mean;
//   k_vel_correct    u -= B flux (flux = -dp/dx from k_flux): the operator is the divergence of
//                    a B flux, so the corrected velocity is divergence-free up to the solve's
//                    residual.
#pragma once

#include "kernels.h"

namespace sydev {

struct VelArgs {
    const double* lo;          // [blk][3] lower corner of every block
    double dlo[3];             // domain lower corner
    double root[3];            // root block edge per direction
    unsigned long long seed;
    int zero_walls;            // 1: velocity 0 on the domain walls
};

__device__ inline unsigned long long splitmix64(unsigned long long x) {
    unsigned long long z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// u [blk][D][NFACE]; one workgroup per block
template <int D, int NB>
__global__ void __launch_bounds__(256) k_face_velocity(VelArgs a, const BlockInfo* __restrict__ binfo, double* u) {
    using G = Geo<D, NB>;
    const int b = blockIdx.x;
    const int lvl = binfo[b].level;
    const double scale = (double)((1ll << (lvl - 1)) * NB);   // cells per root block edge
    long long org[3] = {0, 0, 0};
    for (int d = 0; d < D; d++) org[d] = llrint((a.lo[b * 3 + d] - a.dlo[d]) / (a.root[d] / scale));
    for (int d = 0; d < D; d++) {
        const int fx = G::fx(d), fy = G::fy(d);
        const int ext = d == 2 ? G::NBZ : NB;
        for (int fi = threadIdx.x; fi < G::NFACE; fi += blockDim.x) {
            const int q[3] = {fi % fx, (fi / fx) % fy, fi / (fx * fy)};
            const unsigned long long ii = (unsigned long long)(org[0] + q[0]);
            const unsigned long long jj = (unsigned long long)(org[1] + q[1]);
            const unsigned long long kk = (unsigned long long)(org[2] + q[2]);
            const unsigned long long key = (a.seed * 0x100000001B3ull) ^ ((unsigned long long)lvl << 56) ^
                                           ((unsigned long long)d << 61) ^ (kk << 36) ^ (jj << 18) ^ ii;
            const unsigned long long h = splitmix64(key);
            double v = 2.0 * ((double)(h >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
            if (a.zero_walls) {
                const int s = q[d] == 0 ? 0 : (q[d] == ext ? 1 : -1);
                if (s >= 0 && binfo[b].nbr[2 * d + s] <= PHYS_BNDRY) v = 0.0;
            }
            u[(b * D + d) * G::NFACE + fi] = v;
        }
    }
}

// out [blk][cell] = div (0 on solid cells); part[2 * slot] = (sum dv vf div, sum dv vf) over
// the cells of leaf blocks (0 for other blocks)
template <int D, int NB>
__global__ void __launch_bounds__((Geo<D, NB>::THREADS))
k_divergence(const int* __restrict__ list, const BlockInfo* __restrict__ binfo, const BlockMetric* __restrict__ met,
             const double* __restrict__ af, const double* __restrict__ u, const double* __restrict__ vf,
             double* __restrict__ out, double* __restrict__ part) {
    using G = Geo<D, NB>;
    constexpr int NT = G::THREADS;
    __shared__ double lds[3 * ((NT + 63) / 64)];
    const int b = list[blockIdx.x];
    const BlockMetric mt = met[b];
    const bool leaf = binfo[b].leaf != 0;
    double s0 = 0.0, s1 = 0.0, z = 0.0;
    for (int c = threadIdx.x; c < G::NCELL; c += NT) {
        int i, j, k;
        G::cell_ijk(c, i, j, k);
        double div = 0.0;
        for (int d = 0; d < D; d++) {
            const int lo = (b * D + d) * G::NFACE + G::fidx(d, i, j, k);
            const int hi = (b * D + d) * G::NFACE + G::fidx(d, i + (d == 0), j + (d == 1), k + (d == 2));
            const double fh = (af[hi] * u[hi]) * mt.ds[d];
            const double fl = (af[lo] * u[lo]) * mt.ds[d];
            div += fh - fl;
        }
        const double v = vf[b * G::NCELL + c];
        const double r = v > 0 ? div / (mt.dv * v) : 0.0;
        out[b * G::NCELL + c] = r;
        if (leaf && v > 0) {
            s0 += r * (mt.dv * v);
            s1 += mt.dv * v;
        }
    }
    block_reduce3<NT>(s0, s1, z, lds);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s0;
        part[2 * blockIdx.x + 1] = s1;
    }
}

// u[blk][d][face] -= B[blk][d][face] * flux[blk][face] on the listed blocks
template <int D, int NB>
__global__ void __launch_bounds__(256) k_vel_correct(const int* __restrict__ list, const double* __restrict__ bcoef,
                                                     const double* __restrict__ flux, double* __restrict__ u, int d) {
    using G = Geo<D, NB>;
    const int b = list[blockIdx.x];
    for (int fi = threadIdx.x; fi < G::NFACE; fi += blockDim.x) {
        const int q = (b * D + d) * G::NFACE + fi;
        u[q] = u[q] - bcoef[q] * flux[b * G::NFACE + fi];
    }
}

}  // namespace sydev
