// mixed.h -- FP32 multigrid preconditioner inside the FP64 BiCGSTAB (no reference
// counterpart: the reference runs everything in FP64 with post-smoothing only).
//
// V(l, f_l) approximately solves A_l u_l = f_l on tower T(l) = leaves of levels < l plus
// every block of level l (the reference's MG level, SayakaTower.h:121-129):
//   l == 1 : FP64 BiCGSTAB on the level-1 blocks to a 1e-2 relative residual, with the
//            two-level (Jacobi + one coarse unknown per block) preconditioner of bottom.h;
//   else   : u_l = 0; nu1 red-black sweeps; r = f_l - A_l u_l;
//            f_{l-1} = r on the leaves shared with T(l-1), and the volume-fraction weighted
//            average of r from the children on the non-leaf blocks of level l-1;
//            u_{l-1} = V(l-1, f_{l-1});
//            u_l += u_{l-1} on the shared leaves, += injection of u_{l-1} (zero where vf=0)
//            on the blocks of level l; nu2 red-black sweeps.
// Sweeps, residual, transfer and accumulation are FP32; the operator weights are the FP64
// weights rounded once to FP32.  A sweep (k_sweep32) is one launch over the whole tower,
// double-buffered (u -> t32 -> u): each block runs red then black in LDS, its one-cell halo
// read as it was before the launch (block Gauss-Seidel, Jacobi between blocks), so blocks of
// every sub-level update together and the result does not depend on workgroup order.  The
// first pre-smoothing sweep starts from zero without reading u; the residual launch hands r
// straight to f_{l-1} (copy on the shared leaves, restriction onto the parents), and one
// launch adds the coarse correction.  Each level keeps its own u and f (the towers share
// their coarse leaves, so one in-place array as in the reference cannot hold both
// corrections).  The stale block boundaries take away post-smoothing, so the hybrid cycle runs
// more post-sweeps than asked (one more on 16^3 blocks, three more on 8^3 blocks, where more than
// half the cells sit on a boundary).  Exact red-black colour passes (two launches per sweep)
// remain selectable.
#pragma once

#include "bottom.h"
#include "kernels.h"

struct sayaka_dev_ctx;

namespace sydev {

constexpr int MIX_MAXLV = 16;

struct MixedState {
    int nblk = 0, ncell = 0, coef_version = -1, nlev = 0;
    size_t cap_cells = 0, cap_faces = 0, cap_cf = 0;
    float* w32 = nullptr;
    float* wcf32 = nullptr;
    float* vf32 = nullptr;
    float* t32 = nullptr;       // second buffer of the double-buffered smoother
    // 2: hybrid sweeps (k_sweep32, one launch per sweep over the tower; the default);
    // 1: exact red-black, one launch per colour over the tower, double-buffered;
    // 0: exact red-black, one launch per sub-level and colour in place (the reference's order)
    int merged_smooth = 2;
    float* u[MIX_MAXLV + 1] = {};
    float* f[MIX_MAXLV + 1] = {};
    // x-boundary columns of u[l] and of t32 (hybrid sweeps; load_halo's xslab)
    float* xs[MIX_MAXLV + 1] = {};
    float* xst = nullptr;
    // The reference's bottom test (1e-6 relative + 1e-6 absolute) stops the coarse solve
    // early once the outer residual is small, which leaves the cycle without a coarse
    // correction late in the solve and can stall the outer BiCGSTAB (20 iterations instead of
    // 4 on one C1 seed); a relative-only 1e-2 reduction keeps the cycle's quality constant.
    double bottom_tol_rel = 1.0e-2, bottom_tol_abs = 0.0;
    int bottom_max_iter = 2000;
};

inline void mixed_free(MixedState& ms) {
    float* ptrs[] = {ms.w32, ms.wcf32, ms.vf32, ms.t32, ms.xst};
    for (float* p : ptrs) if (p) (void)hipFree(p);
    ms.xst = nullptr;
    for (int l = 0; l <= MIX_MAXLV; l++) {
        if (ms.u[l]) (void)hipFree(ms.u[l]);
        if (ms.f[l]) (void)hipFree(ms.f[l]);
        if (ms.xs[l]) (void)hipFree(ms.xs[l]);
        ms.xs[l] = nullptr;
        ms.u[l] = ms.f[l] = nullptr;
    }
    ms.w32 = ms.wcf32 = ms.vf32 = ms.t32 = nullptr;
    ms.cap_cells = ms.cap_faces = ms.cap_cf = 0;
    ms.coef_version = -1;
}

__global__ void k_to_f32(const double* __restrict__ src, float* __restrict__ dst, size_t n) {
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        dst[q] = (float)src[q];
}

enum MixOp { M_ZERO = 0, M_F64_TO_F32, M_F32_TO_F64 };

template <int D, int NB>
__global__ void __launch_bounds__((Geo<D, NB>::THREADS))
k_mix(const int* list, int op, const BlockInfo* __restrict__ binfo, const float* __restrict__ vf,
      float* dst, const float* src, double* dst64, const double* src64) {
    using G = Geo<D, NB>;
    const int b = list[blockIdx.x];
    for (int c = threadIdx.x; c < G::NCELL; c += G::THREADS) {
        const int gi = b * G::NCELL + c;
        switch (op) {
            case M_ZERO: dst[gi] = 0.0f; break;
            case M_F64_TO_F32: dst[gi] = (float)src64[gi]; break;
            case M_F32_TO_F64: dst64[gi] = (double)src[gi]; break;
            default: break;
        }
    }
}

// The block's x-boundary columns (i = 0, i = NB-1) from its LDS tile into the slab array
// [block][side][j + NB k] that load_halo reads for x faces.
template <int D, int NB, int NT>
__device__ inline void store_xslab(const float* tile, float* xs, int b) {
    using G = Geo<D, NB>;
    for (int t = threadIdx.x; t < 2 * G::NSLAB; t += NT) {
        const int side = t / G::NSLAB, p = t % G::NSLAB;
        xs[(b * 2 + side) * G::NSLAB + p] = tile[G::tidx(side ? NB - 1 : 0, p % NB, p / NB)];
    }
}

// One hybrid Gauss-Seidel sweep of one block (FP32, x -> out, double-buffered): both colours
// of the block in LDS, red first, then black from the updated red; the one-cell halo holds the
// neighbours' values from before the launch for both colours (block Gauss-Seidel, Jacobi
// between blocks).  A sweep is one launch that reads every operand once, against two
// colour launches that each read every cache line, and the result does not depend on the
// order in which workgroups run.  ZERO: the sweep starts from x = 0 (the first sweep of a
// level's correction), so neither the block nor its halo is read.  RHS64 (with ZERO, the first
// sweep of the cycle): the right-hand side is read in FP64 from rhs64 and stored in FP32 to
// a.out2 for the rest of the cycle.  out64 (optional): the result also in FP64 (the last sweep
// of the cycle writes the preconditioned vector).  Per cell the arithmetic is k_smooth's
// (cell_op).
// Threads of a k_sweep32 workgroup: 16^3 blocks take 1 024 (two updates per thread and colour,
// 62 VGPRs, 8 waves per SIMD); 512 threads held four updates per colour in flight at 112 VGPRs
// (4 waves per SIMD) and ran the 139 M-cell sweep at 3.04 against 3.37 TB/s.
template <int D, int NB>
struct SweepCfg {
    static constexpr int NT = NB >= 16 ? 1024 : Geo<D, NB>::THREADS;
};
template <int D, int NB, bool ZERO, bool RHS64 = false>
__global__ void __launch_bounds__((SweepCfg<D, NB>::NT))
k_sweep32(StencilArgsT<float> a, double* out64, const double* __restrict__ rhs64, const float* __restrict__ xsin,
          float* xsout) {
    using G = Geo<D, NB>;
    using T = float;
    static_assert(ZERO || !RHS64, "the FP64 right-hand side is read by the first sweep");
    // cell inputs without the right-hand side when it comes from rhs64
    constexpr int LMODE = RHS64 ? ST_APPLY : ST_SMOOTH;
    auto load_in = [&](CellIn<T>& ci, int b, int c) {
        load_cell<D, NB, LMODE, T>(ci, a, b, c, true);
        if (RHS64) {
            ci.rhs = (T)rhs64[b * G::NCELL + c];
            a.out2[b * G::NCELL + c] = ci.rhs;
        }
    };
    static_assert(NB % 2 == 0, "even block edge");
    constexpr int NT = SweepCfg<D, NB>::NT;
    constexpr int CPT = (G::NCELL + NT - 1) / NT;
    constexpr int NACT = G::NCELL / 2;
    constexpr int APT = (NACT + NT - 1) / NT;
    constexpr bool PRE = APT == 1;   // one update per thread and colour: inputs loaded at entry
    __shared__ __attribute__((aligned(16))) T tile[G::NTILE];
    __shared__ int biw[BI_WORDS];
    const int b = a.list[blockIdx.x];
    if ((int)threadIdx.x < BI_WORDS) biw[threadIdx.x] = reinterpret_cast<const int*>(a.binfo + b)[threadIdx.x];
    T xo[CPT];
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int c = threadIdx.x + q * NT;
        xo[q] = (!ZERO && c < G::NCELL) ? a.x[b * G::NCELL + c] : T(0);
    }
    CellIn<T> in[2];
    if (PRE) {
        const int ia = threadIdx.x;
#pragma unroll
        for (int col = 0; col < 2; col++) {
            if (ia < NACT) {
                int i, j, k;
                colour_cell<NB>(ia, col, i, j, k);
                load_in(in[col], b, G::cidx(i, j, k));
            }
        }
    }
    const T dv = (T)a.met[b].dv;
    __syncthreads();
    const BlockInfo& bi = *reinterpret_cast<const BlockInfo*>(biw);
    if (ZERO) {
        for (int t = threadIdx.x; t < G::NTILE; t += NT) tile[t] = T(0);
    } else {
#pragma unroll
        for (int q = 0; q < CPT; q++) {
            const int c = threadIdx.x + q * NT;
            if (c >= G::NCELL) continue;
            int i, j, k;
            G::cell_ijk(c, i, j, k);
            tile[G::tidx(i, j, k)] = xo[q];
        }
        load_halo<D, NB, T, NT>(tile, a.x, a.w, bi, a.m, xsin);
    }
    if (PRE) {
        const int ia = threadIdx.x;
#pragma unroll
        for (int col = 0; col < 2; col++) {
            if (ia < NACT) {
                int i, j, k;
                colour_cell<NB>(ia, col, i, j, k);
                load_wcf<D, NB, T>(in[col], a, b, G::cidx(i, j, k), bi);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int col = 0; col < 2; col++) {
        // a cell of one colour reads only cells of the other colour (and itself), so the
        // colour's updates go into the tile without a barrier between reads and writes
#pragma unroll 1
        for (int q = 0; q < APT; q++) {
            const int ia = threadIdx.x + q * NT;
            if (ia >= NACT) continue;
            int i, j, k;
            colour_cell<NB>(ia, col, i, j, k);
            CellIn<T> tmp;
            CellIn<T>& ci = PRE ? in[col] : tmp;
            if (!PRE) {
                load_in(ci, b, G::cidx(i, j, k));
                load_wcf<D, NB, T>(ci, a, b, G::cidx(i, j, k), bi);
            }
            const T u = cell_op<D, NB, true, T>(tile, ci.w, ci.wcf, i, j, k, bi, a.m, a.alpha, dv, ci.vf, ci.rhs);
            tile[G::tidx(i, j, k)] = u;
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int c = threadIdx.x + q * NT;
        if (c >= G::NCELL) continue;
        int i, j, k;
        G::cell_ijk(c, i, j, k);
        const T v = tile[G::tidx(i, j, k)];
        a.out[b * G::NCELL + c] = v;
        if (out64) out64[b * G::NCELL + c] = (double)v;
    }
    if (xsout) store_xslab<D, NB, NT>(tile, xsout, b);
}

// r = f - A u on tower T(l) (FP32), handed straight to level l-1's right-hand side: a block
// below level l (a leaf that T(l-1) shares) passes r on as it is; a block of level l restricts
// r onto its parent's octant: the vf-weighted average of the 2^d children's r (sum of r * vf
// in kk, jj, ii order, / 2^d, / the parent's vf; 0 where that is 0).  Replaces the residual launch, the copy and the restriction (and their
// re-reads of r).  a.x = u, a.rhs = f; fc = f_{l-1}.
template <int D, int NB>
__global__ void __launch_bounds__((Geo<D, NB>::THREADS))
k_resid_restrict32(StencilArgsT<float> a, float* fc, int l, const float* __restrict__ xsin) {
    using G = Geo<D, NB>;
    using T = float;
    constexpr int NT = G::THREADS;
    constexpr int CPT = (G::NCELL + NT - 1) / NT;
    constexpr bool PRE = CPT == 1;
    __shared__ __attribute__((aligned(16))) T tile[G::NTILE];
    __shared__ int biw[BI_WORDS];
    const int b = a.list[blockIdx.x];
    if ((int)threadIdx.x < BI_WORDS) biw[threadIdx.x] = reinterpret_cast<const int*>(a.binfo + b)[threadIdx.x];
    CellIn<T> in[1];
    T xo[CPT];
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int c = threadIdx.x + q * NT;
        if (c >= G::NCELL) continue;
        if (PRE) {
            load_cell<D, NB, ST_RESID, T>(in[0], a, b, c, true);
            xo[q] = in[0].x;
        } else {
            xo[q] = a.x[b * G::NCELL + c];
        }
    }
    const T dv = (T)a.met[b].dv;
    __syncthreads();
    const BlockInfo& bi = *reinterpret_cast<const BlockInfo*>(biw);
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int c = threadIdx.x + q * NT;
        if (c >= G::NCELL) continue;
        int i, j, k;
        G::cell_ijk(c, i, j, k);
        tile[G::tidx(i, j, k)] = xo[q];
    }
    load_halo<D, NB, T>(tile, a.x, a.w, bi, a.m, xsin);
    if (PRE && (int)threadIdx.x < G::NCELL) load_wcf<D, NB, T>(in[0], a, b, threadIdx.x, bi);
    __syncthreads();
    const bool below = bi.level < l;   // uniform over the workgroup
    T rv[CPT];
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        rv[q] = T(0);
        const int c = threadIdx.x + q * NT;
        if (c >= G::NCELL) continue;
        int i, j, k;
        G::cell_ijk(c, i, j, k);
        if (!PRE) {
            load_cell<D, NB, ST_RESID, T>(in[0], a, b, c, true);
            load_wcf<D, NB, T>(in[0], a, b, c, bi);
        }
        const T vfv = in[0].vf;
        const T Lp = cell_op<D, NB, false, T>(tile, in[0].w, in[0].wcf, i, j, k, bi, a.m, a.alpha, dv, vfv, T(0));
        const T r = vfv > 0 ? in[0].rhs - Lp : T(0);
        if (below) fc[b * G::NCELL + c] = r;
        else rv[q] = r * vfv;
    }
    if (below) return;
    __syncthreads();   // every halo read of the tile is done: it now holds r * vf by cell index
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int c = threadIdx.x + q * NT;
        if (c < G::NCELL) tile[c] = rv[q];
    }
    __syncthreads();
    constexpr int H = NB / 2;
    constexpr int NPC = G::NCELL / G::NCHILD;   // parent cells covered by this block
    const int p = bi.parent, wc = bi.which_child;
    const T wgt = 1.0f / G::NCHILD;
    for (int t = threadIdx.x; t < NPC; t += NT) {
        const int pi = t % H, pj = (t / H) % H, pk = D == 3 ? t / (H * H) : 0;
        T sum = T(0);
        for (int kk = 0; kk <= (D == 3 ? 1 : 0); kk++)
            for (int jj = 0; jj <= 1; jj++)
                for (int ii = 0; ii <= 1; ii++) sum += tile[G::cidx(2 * pi + ii, 2 * pj + jj, 2 * pk + kk)];
        const int pc = G::cidx(pi + side_of(wc, 0) * H, pj + side_of(wc, 1) * H, D == 3 ? pk + side_of(wc, 2) * H : 0);
        const T v = a.vf[p * G::NCELL + pc];
        fc[p * G::NCELL + pc] = v > 0.0f ? (sum * wgt) / v : 0.0f;
    }
}

// The coarse correction on tower T(l): out = u_l + u_{l-1} on the blocks below level l (the
// leaves T(l-1) shares), u_l + the injection of the parent's u_{l-1} (nothing where vf = 0) on
// the blocks of level l, in one launch over T(l).  out may be
// u (in place) or the other buffer of the double-buffered sweeps that follow.
template <int D, int NB>
__global__ void __launch_bounds__((Geo<D, NB>::THREADS))
k_correct32(const int* list, const BlockInfo* __restrict__ binfo, const float* __restrict__ vf, const float* u,
            const float* __restrict__ uc, float* out, int l, float* xsout) {
    using G = Geo<D, NB>;
    const int b = list[blockIdx.x];
    const BlockInfo& bi = binfo[b];
    const bool below = bi.level < l;
    const int h = NB / 2;
    const int p = bi.parent, wc = bi.which_child;
    for (int c = threadIdx.x; c < G::NCELL; c += G::THREADS) {
        const int gi = b * G::NCELL + c;
        float v = u[gi];
        if (below) {
            v += uc[gi];
        } else {
            int i, j, k;
            G::cell_ijk(c, i, j, k);
            const int pc = G::cidx((i >> 1) + side_of(wc, 0) * h, (j >> 1) + side_of(wc, 1) * h,
                                   D == 3 ? (k >> 1) + side_of(wc, 2) * h : 0);
            if (vf[gi] > 0.0f) v += uc[p * G::NCELL + pc];
        }
        out[gi] = v;
        const int i = c % NB;
        if (xsout && (i == 0 || i == NB - 1)) xsout[(b * 2 + (i ? 1 : 0)) * G::NSLAB + c / NB] = v;
    }
}

}  // namespace sydev
