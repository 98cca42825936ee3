// bottom.h -- coarsest-level solve as one cooperative persistent kernel.
//
// The reference solves the level-1 problem with unpreconditioned BiCGSTAB to 1e-6
// relative + 1e-6 absolute Linf residual, up to 2000 iterations, twice per outer
// iteration (mg_solve_bottom -> cg_solve_level(1, ...), SayakaMacSolver.cpp:164-242,
// 683-920).  Host round trips per iteration would dominate, so the whole loop runs in one
// launch: one workgroup per level-1 block, every Krylov vector in registers, the two
// vectors the operator reads across blocks (p and s) mirrored to global memory, and every
// scalar reduced into all workgroups through grid barriers, so each workgroup evaluates the
// same convergence and breakdown tests and leaves the loop at the same iteration.
#pragma once

#include <hip/hip_cooperative_groups.h>

#include "kernels.h"

namespace sydev {

struct BottomArgs {
    StencilArgs st;        // list = level-1 blocks, m = 1
    const double* rhs;     // RESID
    double* sol;           // CORR (zeroed, written on convergence)
    double* pg;            // mirror of p  [blk][NCELL]
    double* sg;            // mirror of s
    double* partials;      // 2 buffers x 3 values x gridDim.x
    int* stats;            // [0] calls, [1] total iterations, [2] last iterations, [3] last status
    double tol_rel, tol_abs;
    int max_iter;
    int exact;             // reference-order sums (see RedBuf::exact)
};

// Two sums and a max over every cell of every workgroup, identical in all workgroups.
template <int NT, int NCELL, int CPT>
__device__ inline void coop_reduce(cooperative_groups::grid_group& grid, const double (&q0)[CPT],
                                   const double (&q1)[CPT], int nsum, double& s0, double& s1, double& mx,
                                   double* part, double* scratch, double* chunk, double* lds, int exact) {
    double a0 = 0.0, a1 = 0.0;
    if (exact) {
        if (nsum > 0) a0 = seq_block_sum<NT, NCELL, CPT>(q0, scratch);
        if (nsum > 1) a1 = seq_block_sum<NT, NCELL, CPT>(q1, scratch);
        double z0 = 0.0, z1 = 0.0;
        block_reduce3<NT>(z0, z1, mx, lds);
    } else {
#pragma unroll
        for (int q = 0; q < CPT; q++) {
            a0 += q0[q];
            a1 += q1[q];
        }
        block_reduce3<NT>(a0, a1, mx, lds);
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = a0;
        part[3 * blockIdx.x + 1] = a1;
        part[3 * blockIdx.x + 2] = mx;
    }
    grid.sync();
    double ga, gb, gm;
    sum_published<NT>(part, gridDim.x, exact != 0, ga, gb, gm, chunk, lds);
    __shared__ double bc[3];
    if (threadIdx.x == 0) {
        bc[0] = ga;
        bc[1] = gb;
        bc[2] = gm;
    }
    __syncthreads();
    s0 = bc[0];
    s1 = bc[1];
    mx = bc[2];
    __syncthreads();
}

template <int D, int NB>
__global__ void __launch_bounds__((Geo<D, NB>::THREADS)) k_bottom(BottomArgs A) {
    using G = Geo<D, NB>;
    constexpr int NT = G::THREADS;
    constexpr int CPT = (G::NCELL + NT - 1) / NT;
    namespace cg = cooperative_groups;
    cg::grid_group grid = cg::this_grid();
    __shared__ double tile[G::NTILE];
    __shared__ double lds[3 * ((NT + 63) / 64)];
    __shared__ double chunk[3 * NT];
    const StencilArgs& a = A.st;
    const int b = a.list[blockIdx.x];
    __shared__ BlockInfo bi;
    if (threadIdx.x == 0) bi = a.binfo[b];
    __syncthreads();
    const double dv = a.met[b].dv;
    double* partA = A.partials;
    double* partB = A.partials + 3 * gridDim.x;
    int pb = 0;
    auto part = [&]() { double* p = pb ? partB : partA; pb ^= 1; return p; };

    double x[CPT], r[CPT], rh[CPT], p[CPT], v[CPT], s[CPT], t[CPT], vfv[CPT], q0[CPT], q1[CPT];
    // x0 = 0 (CORR zeroed by mg_solve_bottom), r = rhs - A*0 = rhs on fluid cells
    double s0 = 0.0, s1 = 0.0, mx = 0.0;
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int c = threadIdx.x + q * NT;
        x[q] = r[q] = rh[q] = p[q] = v[q] = s[q] = t[q] = q0[q] = q1[q] = 0.0;
        vfv[q] = 0.0;
        if (c < G::NCELL) {
            const size_t gi = (size_t)b * G::NCELL + c;
            A.sol[gi] = 0.0;
            vfv[q] = a.vf[gi];
            r[q] = vfv[q] > 0 ? A.rhs[gi] - 0.0 : 0.0;
            rh[q] = r[q];
            q0[q] = rh[q] * r[q];
            mx = fmax(mx, fabs(r[q]));
        }
    }
    coop_reduce<NT, G::NCELL, CPT>(grid, q0, q1, 1, s0, s1, mx, part(), tile, chunk, lds, A.exact);
    const double rnorm0 = mx;
    double rho = s0, rho1 = 0.0, alpha = 0.0, omega = 0.0;
    int iter = 0, conv = 0, stat = 0;
    if (rnorm0 != 0.0) {
        for (iter = 1; iter <= A.max_iter; iter++) {
            if (rho == 0.0) { stat = 1; break; }
            const double beta = iter == 1 ? 0.0 : (rho / rho1) * (alpha / omega);
#pragma unroll
            for (int q = 0; q < CPT; q++) {
                const int c = threadIdx.x + q * NT;
                if (iter == 1) p[q] = r[q];
                else {
                    const double p1 = 1.0 * p[q] + -omega * v[q];
                    p[q] = 1.0 * r[q] + beta * p1;
                }
                if (c < G::NCELL) A.pg[(size_t)b * G::NCELL + c] = p[q];
            }
            grid.sync();
            load_tile<D, NB, double>(tile, A.pg, bi, b);
            __syncthreads();
            mx = 0.0;
#pragma unroll
            for (int q = 0; q < CPT; q++) {
                const int c = threadIdx.x + q * NT;
                q0[q] = 0.0;
                if (c < G::NCELL) {
                    int i, j, k;
                    G::cell_ijk(c, i, j, k);
                    v[q] = cell_op<D, NB, false, double>(tile, b, i, j, k, bi, 1, a, dv, vfv[q], 0.0);
                    q0[q] = rh[q] * v[q];
                }
            }
            coop_reduce<NT, G::NCELL, CPT>(grid, q0, q1, 1, s0, s1, mx, part(), tile, chunk, lds, A.exact);
            if (s0 == 0.0) { stat = 1; break; }
            alpha = rho / s0;
            mx = 0.0;
#pragma unroll
            for (int q = 0; q < CPT; q++) {
                const int c = threadIdx.x + q * NT;
                s[q] = 1.0 * r[q] + -alpha * v[q];
                x[q] = 1.0 * x[q] + alpha * p[q];
                q0[q] = 0.0;
                if (c < G::NCELL) {
                    A.sg[(size_t)b * G::NCELL + c] = s[q];
                    mx = fmax(mx, fabs(s[q]));
                }
            }
            coop_reduce<NT, G::NCELL, CPT>(grid, q0, q1, 0, s0, s1, mx, part(), tile, chunk, lds, A.exact);
            if (mx < rnorm0 * A.tol_rel || mx < A.tol_abs) { conv = 1; break; }
            load_tile<D, NB, double>(tile, A.sg, bi, b);
            __syncthreads();
            mx = 0.0;
#pragma unroll
            for (int q = 0; q < CPT; q++) {
                const int c = threadIdx.x + q * NT;
                q0[q] = q1[q] = 0.0;
                if (c < G::NCELL) {
                    int i, j, k;
                    G::cell_ijk(c, i, j, k);
                    t[q] = cell_op<D, NB, false, double>(tile, b, i, j, k, bi, 1, a, dv, vfv[q], 0.0);
                    q0[q] = t[q] * s[q];
                    q1[q] = t[q] * t[q];
                }
            }
            double ts = 0.0, tt = 0.0;
            coop_reduce<NT, G::NCELL, CPT>(grid, q0, q1, 2, ts, tt, mx, part(), tile, chunk, lds, A.exact);
            if (tt == 0.0) { stat = 1; break; }
            omega = ts / tt;
            mx = 0.0;
#pragma unroll
            for (int q = 0; q < CPT; q++) {
                const int c = threadIdx.x + q * NT;
                x[q] = 1.0 * x[q] + omega * s[q];
                r[q] = 1.0 * s[q] + -omega * t[q];
                q0[q] = 0.0;
                if (c < G::NCELL) {
                    mx = fmax(mx, fabs(r[q]));
                    q0[q] = rh[q] * r[q];
                }
            }
            coop_reduce<NT, G::NCELL, CPT>(grid, q0, q1, 1, s0, s1, mx, part(), tile, chunk, lds, A.exact);
            if (mx < rnorm0 * A.tol_rel + A.tol_abs) { conv = 1; break; }
            if (omega == 0.0) { stat = 1; break; }
            rho1 = rho;
            rho = s0;
        }
        if (iter > A.max_iter) iter = A.max_iter;
    }
    if (conv) {
#pragma unroll
        for (int q = 0; q < CPT; q++) {
            const int c = threadIdx.x + q * NT;
            if (c < G::NCELL) A.sol[(size_t)b * G::NCELL + c] = 0.0 + x[q];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        A.stats[0] += 1;
        A.stats[1] += iter;
        A.stats[2] = iter;
        A.stats[3] = stat;
    }
}

}  // namespace sydev
