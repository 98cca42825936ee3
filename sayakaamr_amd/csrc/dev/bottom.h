// bottom.h -- coarsest-level solve as one persistent kernel with its own grid barrier.
//
// The reference solves the level-1 problem with unpreconditioned BiCGSTAB to 1e-6
// relative + 1e-6 absolute Linf residual, up to 2000 iterations, twice per outer
// iteration (mg_solve_bottom -> cg_solve_level(1, ...), SayakaMacSolver.cpp:164-242,
// 683-920).  Host round trips per iteration would dominate, so the whole loop runs in one
// launch.  Grid barriers are the cost that remains (about 1.7 us at 8 workgroups, 9.5 us at
// 64 on MI355X), so the kernel is shaped to need few of them:
//
//  * BPW level-1 blocks per workgroup (4 for 8^3 blocks: 16 workgroups for a 4^3 root, the
//    most blocks whose Krylov vectors still fit in registers), every vector in registers;
//  * three barriers per iteration instead of five.  The operator needs p and s on the
//    ghost cells (the neighbours' boundary cells).  Instead of publishing p and s and
//    waiting for them, each workgroup recomputes them on its ghost cells from the values it
//    has read anyway -- r and v are published at the barriers that carry the dot products,
//    and p on the ghosts is kept from the previous iteration -- with the same expression
//    and the same operands as the owning workgroup, so the ghost copies are bit-identical;
//  * the half-step test on |s| travels with the t.s / t.t sums (t is computed before the
//    test and discarded when it passes; x and the iteration count are unchanged).
//
// Reductions: exact mode reproduces the reference's order (DotProd, SayakaBlockData.cpp:
// 185-208: cells k,j,i of a block in sequence; mg_dotprod_level, SayakaMacUtil.cpp:79-106:
// block sums in tower order in sequence).  The per-block chains run concurrently in lanes
// of one wave (one lane per block and sum), the chain over blocks after the barrier.
//
// Two-level variant (PC = true, used by the mixed-precision cycle, never by the reference
// path): right-preconditioned BiCGSTAB with
//     M^-1 x = x / diag(A) + P S0^-1 P^T D x,
// D = dv * vf per cell, P the level-1-block indicator (one coarse unknown per block, zero on
// solid cells), S0 = P^T (D A) P the block-to-block coupling (D A is the symmetric
// weighted graph Laplacian of the face weights).  The coarse right-hand sides P^T D p and
// P^T D s follow from per-block sums of D r and D v, which travel with the barriers that
// already carry the dot products (p = r + beta (p - omega v), s = r - alpha v), and the
// preconditioned values on ghost cells are x / diag(A) there plus the neighbour block's
// coarse value, so the preconditioner costs no extra barrier.  On the 3D solve's level-1
// problem (32^3 cells, density ratio 1000) it cuts BiCGSTAB from ~50 to ~15 iterations for a
// 1e-2 residual reduction.
#pragma once

#include "kernels.h"

namespace sydev {

// Phase timing (the library built with SAYAKA_BOTTOM_PHASES, lib/libsayaka_dev_phases.so; the
// product build compiles the marks to nothing): thread 0 of workgroup 0 reads the wall clock at
// the phase boundaries of every iteration -- after its workgroup barriers, so the phase ends
// for the whole workgroup -- and adds the durations to g_bottom_phases at the end of a launch
// (index = BottomPhase; [30] iterations, [31] launches).
enum BottomPhase {
    PH_OP_V = 0,            // p update, operand tile, v = A p, staging of rhat . v
    PH_A = 1,               // reduction A: chains, grid barrier, totals (1, 2, 3)
    PH_OP_T = 4,            // v ghosts, s / x update, operand tile, t = A s, staging of t.s, t.t
    PH_B = 5,               // reduction B (5, 6, 7)
    PH_UPD_R = 8,           // x / r update, staging of rhat . r
    PH_C = 9,               // reduction C (9, 10, 11)
    PH_R_GHOST = 12,        // r ghosts
    PH_INIT = 13,           // the launch's first reduction (13, 14, 15)
    PH_WAIT = 16,           // per reduction: the staging wait before its chains (16 + reduction)
    PH_NUM = 20
};
#ifdef SAYAKA_BOTTOM_PHASES
__device__ unsigned long long g_bottom_phases[32];
__device__ inline void bottom_phase_mark(unsigned long long* pht, int k) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const unsigned long long t = wall_clock64();
        pht[1 + k] += t - pht[0];
        pht[0] = t;
    }
}
#define SY_PH_MARK(pht, k) bottom_phase_mark(pht, k)
#else
#define SY_PH_MARK(pht, k) ((void)(pht), (void)(k))
#endif

struct BottomArgs {
    StencilArgs st;        // list = level-1 blocks (tower 1), nlist of them, m = 1
    const double* rhs;     // RESID
    double* sol;           // CORR (written: zero, then the solution on convergence)
    double* vpub;          // published v  [blk][NCELL]
    double* rpub;          // published r
    double* partials;      // 3 buffers x 3 values x nlist
    int* stats;            // [0] calls, [1] total iterations, [2] last iterations, [3] last status,
                           // [4] grid-barrier time-outs
    unsigned long long* flags;     // per workgroup: (last barrier episode it reached) << 4 | its XCD
    unsigned long long episode0;   // episodes of this launch are above every earlier launch's
    int confine;           // grid of 8 x the workgroups: the workers are the workgroups blockIdx.x % 8 == 0
    double tol_rel, tol_abs;
    int max_iter;
    int exact;             // reference-order sums (see RedBuf::exact)
    // two-level preconditioner (k_bottom<.., true>): nlist <= PC_MAXBLK
    const float* s0inv;    // coarse inverse, row-major nlist x nlist
    const double* jac;     // [level-1 list position][cell]: 1 / diag(A), 0 on solid cells
    const int* l1pos;      // block id -> position in the level-1 list, -1 for other blocks
    const int* skip;       // when non-null and *skip != 0: the launch does nothing (a solve that
                           // stopped at the half step queued this V-cycle before knowing it)
};

constexpr int PC_MAXBLK = 64;   // level-1 blocks the two-level preconditioner supports

// The XCD (accelerator complex die) the calling wave runs on: hardware register XCC_ID, bits 3:0.
__device__ inline unsigned xcd_id() { return __builtin_amdgcn_s_getreg((3 << 11) | 20) & 15u; }

// Grid-barrier state of one workgroup of the persistent launch.
struct BarState {
    unsigned long long ep;   // episodes reached so far
    int wid;                 // worker index: flag slot, list offset / BPW
    int* broken;             // (LDS) a barrier of this workgroup timed out
    int* light;              // (LDS) every worker is on one XCD: barriers without L2 maintenance
};
constexpr int BOTTOM_STALL = 25;   // PC: iterations without halving the best residual before stopping
constexpr unsigned BARRIER_SPIN_MAX = 1u << 18;   // polls before a grid barrier gives up (~0.2 s)

template <int D, int NB>
struct BotCfg {
    using G = Geo<D, NB>;
    static constexpr int BPW = G::NCELL >= 4096 ? 1 : 4;              // blocks per workgroup
    static constexpr int WCELL = BPW * G::NCELL;
    static constexpr int NT = WCELL < 512 ? WCELL : 512;
    static constexpr int CPT = WCELL / NT;                             // cells per thread
    static constexpr int WGHOST = BPW * G::NGHOST;
    static constexpr int GPT = (WGHOST + NT - 1) / NT;                 // ghost slots per thread
    // staged terms of the ordered sums, chain l = which * BPW + blk.  512-term chains
    // (seg_chain512_asm: lane j reads terms 16 j .. 16 j + 15) put two empty slots after every
    // 16 terms, so that the lanes' 128-byte runs start 144 bytes apart and spread over the LDS
    // banks; other lengths are contiguous (lane_chain)
    static constexpr bool SEGCH = G::NCELL == 512;
    static constexpr int CSTRIDE = SEGCH ? G::NCELL + G::NCELL / 8 + 2 : G::NCELL + 2;
    __host__ __device__ static constexpr int chain_slot(int l, int c) {
        return l * CSTRIDE + (SEGCH ? c + 2 * (c >> 4) : c);
    }
    static constexpr int NRED = 2 * BPW * CSTRIDE;                     // two sums staged
    static constexpr int NTIL = BPW * G::NTILE;
    // face weights of the workgroup's blocks staged in LDS when they fit beside the rest
    static constexpr bool WLDS = BPW * D * G::NFACE * 8 <= 64 * 1024;
    static constexpr int NWSH = WLDS ? BPW * D * G::NFACE : 1;
    static constexpr int NSHARED = NRED > NTIL ? (NRED > 3 * NT ? NRED : 3 * NT) : (NTIL > 3 * NT ? NTIL : 3 * NT);
    static constexpr int NSHARED_FAST = NTIL > 3 * NT ? NTIL : 3 * NT;   // no ordered chains
    static constexpr int NW = NT / 64;                                    // waves per workgroup
    static_assert(WCELL % NT == 0, "cells per workgroup must be a multiple of the thread count");
    static_assert((G::NCELL & (G::NCELL - 1)) == 0, "block cell count must be a power of two");
};

// Operator on one level-1 cell with the wall treatment folded into the staged data: on
// Neumann wall faces the staged weight is 0, on Dirichlet / SimpleFill walls the ghost holds
// -phi_c, so every face contributes off = w and offval = w * ghost.  Term by term (face
// order, zeros added where the reference adds zeros, w * (-phi_c) == -w * phi_c) this is the
// reference's apply (SayakaMacOperator.cpp:1409-1454) with m = 1, where every face is
// same-level or a wall.
template <int D, int NB>
__device__ inline double bot_apply_cell(const double* tile, const double* wb, int i, int j, int k, double alpha,
                                        double dv, double vfv) {
    using G = Geo<D, NB>;
    const double phic = tile[G::tidx(i, j, k)];
    double diag = alpha * dv, offd = 0.0;
#pragma unroll
    for (int f = 0; f < G::NF; f++) {
        const int d = f >> 1;
        const int fi = G::fidx(d, i + (f == 1), j + (f == 3), k + (f == 5));
        const int ni = G::tidx(i + (f == 1) - (f == 0), j + (f == 3) - (f == 2), k + (f == 5) - (f == 4));
        const double off = wb[d * G::NFACE + fi];
        const double offval = off * tile[ni];
        diag += off;
        offd += offval;
    }
    const double Lphi = (phic * diag - offd) / dv;
    return vfv > 0 ? Lphi / vfv : 0.0;
}

template <int D, int NB, bool PC>
struct Bottom {
    using G = Geo<D, NB>;
    using C = BotCfg<D, NB>;
    static constexpr int BPW = C::BPW, NT = C::NT, CPT = C::CPT, GPT = C::GPT, NW = C::NW;
    static constexpr int PS = PC ? 4 : 3;   // values per block in the published partials

    // Per-block sums of the thread's per-block terms bq (two-level variant), published in
    // slot 3 of each block's partials before the barrier that follows.
    __device__ static void publish_block_sums(const double (&bq)[BPW], double* part, double* wq, int w0, int nvalid) {
        __syncthreads();
#pragma unroll
        for (int b = 0; b < BPW; b++) {
            const double v = wave_sum(bq[b]);
            if ((threadIdx.x & 63) == 0) wq[b * NW + (threadIdx.x >> 6)] = v;
        }
        __syncthreads();
        if ((int)threadIdx.x < BPW && (int)threadIdx.x < nvalid) {
            double t = 0.0;
            for (int w = 0; w < NW; w++) t += wq[threadIdx.x * NW + w];
            part[PS * (w0 + threadIdx.x) + 3] = t;
        }
    }
    // After the barrier: every block's published sum into LDS.
    __device__ static void read_block_sums(const double* part, double* sx, int nl) {
        for (int t = threadIdx.x; t < nl; t += NT) sx[t] = load_published(part + PS * t + 3);
        __syncthreads();
    }
    // cx = S0^-1 sx for the nl coarse unknowns: eight lanes per row, combined by shuffles.
    __device__ static void coarse_mv(const float* s0, const double* sx, double* cx, int nl) {
        for (int row0 = 0; row0 < nl; row0 += NT / 8) {
            const int row = row0 + (int)(threadIdx.x >> 3);
            double acc = 0.0;
            if (row < nl)
                for (int j = threadIdx.x & 7; j < nl; j += 8) acc += (double)s0[row * (PC_MAXBLK + 1) + j] * sx[j];
            acc += __shfl_xor(acc, 1, 64);
            acc += __shfl_xor(acc, 2, 64);
            acc += __shfl_xor(acc, 4, 64);
            if (row < nl && (threadIdx.x & 7) == 0) cx[row] = acc;
        }
        __syncthreads();
    }

    // Stage one per-cell term of sum `which` (0, 1) for block blk, cell c.  Exact mode keeps
    // it in LDS for the ordered chains; fast mode adds it to the thread's running sum.
    __device__ static void stage(double* red, int exact, int which, int blk, int c, double val, double& acc) {
        if (exact) {
            red[C::chain_slot(which * BPW + blk, c)] = val;
        } else {
            acc += val;
        }
    }

    // Grid barrier of the persistent launch as a flag gather: each workgroup raises its own flag
    // to the episode number (with its XCD in the low 4 bits) and wave 0 waits on every flag at
    // once, one lane per workgroup: no argument loads and no returning atomic on the critical
    // path.  The wait is bounded: a workgroup that times out counts it in stats[4], and waits no
    // more in this launch (*broken), so every wave ends soon; the host reports the error.
    //
    // Memory ordering.  Device-wide (the grid spread over the XCDs, whose L2 caches are not
    // coherent with each other): the operations of grid.sync -- thread 0 makes the workgroup's
    // stores visible device-wide with an agent-scope release (an L2 write-back), raises the
    // flag, and after the wait an agent-scope acquire invalidates the caches.  One XCD (*light,
    // decided at the launch's first barrier from the flags' XCD bits, which every worker reads
    // alike): the XCD's L2 is the coherence point of every worker, so the release is each
    // wave's wait for its own stores to complete there, and the workers read one another's data
    // with loads that bypass the per-CU L1 (load_published); neither the write-back nor the
    // invalidation is needed.  On MI355X the pair costs 1.5 us per episode against 3.3 us
    // device-wide (16 workgroups, 16 KB exchanged each).
    __device__ static void grid_barrier(const BottomArgs& A, BarState& bs, int nwg) {
        const unsigned long long ep = ++bs.ep;
        const bool first = ep == A.episode0 + 1;
        const int lt = *bs.light;
        if (lt) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_s_waitcnt(0);   // this wave's stores have reached the L2
        }
        __syncthreads();
        const unsigned me = xcd_id();
        if (threadIdx.x == 0) {
            if (!lt) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_store(A.flags + bs.wid, (ep << 4) | me, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (threadIdx.x < 64) {
            int late = 0, other = 0;
            const int skip = *bs.broken;
            for (int w = threadIdx.x; w < nwg && !skip; w += 64) {
                unsigned spins = 0;
                unsigned long long f;
                while (((f = __hip_atomic_load(A.flags + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >> 4) < ep) {
                    if (++spins > BARRIER_SPIN_MAX) {
                        late = 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                other |= (unsigned)(f & 15u) != me;
            }
            if (late) {
                atomicAdd(A.stats + 4, 1);
                *bs.broken = 1;
            }
            if (first) {
                const bool one = __ballot(other || late) == 0;
                if (threadIdx.x == 0) *bs.light = A.confine && one && !skip;
            }
            if (!lt) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            else __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        }
        __syncthreads();
    }

    // Per-block sums of the staged terms (nsum of them) for the blocks of this workgroup,
    // published with the workgroup's max; a grid barrier; then every workgroup reads all
    // partials and forms the totals (in block order when exact).  a0/a1 are the thread's
    // running sums in fast mode.  Both sides of the barrier run in wave 0 alone after one
    // workgroup barrier, so the other waves only wait once on each side.
    struct NoWork {
        __device__ void operator()() const {}
    };
    // after_barrier(): run by every thread right after the grid barrier, beside wave 0's totals
    // (loads of data the other workers published before the barrier, which need no totals)
    template <class AfterBarrier = NoWork>
    __device__ static void reduce(unsigned long long* pht, int ph, BarState& bs,
                                  const BottomArgs& A, double a0, double a1,
                                  int nsum, double mxl, double* part, double* red, double* wl, int nvalid,
                                  double& s0, double& s1, double& mx, const AfterBarrier& after_barrier = AfterBarrier()) {
        const int w0 = bs.wid * BPW;
        const int nl = A.st.nlist;
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        mxl = wave_max(mxl);
        if (lane == 0) wl[wv] = mxl;
        if (PC || !A.exact) {
            a0 = wave_sum(a0);
            a1 = wave_sum(a1);
            if (lane == 0) {
                wl[16 + wv] = a0;
                wl[32 + wv] = a1;
            }
        }
        __syncthreads();
        SY_PH_MARK(pht, PH_WAIT + (ph - 1) / 4);
        // ordered per-block chains (the two-level variant always sums in tree order); chain
        // l = which * BPW + blk is the sum `which` of block blk
        constexpr bool SEGCH = C::SEGCH;
        if constexpr (SEGCH) {
            // two chains per wave with every value in registers (seg_chain512_asm): chain l in
            // half l % 2 of wave l / 2, its result in that half's last lane
            const int nch = BPW * nsum;
            if (!PC && A.exact && 2 * wv < nch) {
                const int l = 2 * wv + (lane >> 5);
                const int lc = l < nch ? l : nch - 1;
                const double sv = seg_chain512_asm<200>(lds_addr(red + C::chain_slot(lc, 16 * (lane & 31))), 0.0);
                if ((lane & 31) == 31 && l < nch && l % BPW < nvalid) part[PS * (w0 + l % BPW) + l / BPW] = sv;
            }
        }
        if (wv == 0) {
            if (!PC && A.exact) {
                // one lane per chain
                if (!SEGCH && lane < BPW * nsum) {
                    const double sv = lane_chain<G::NCELL, 200>(red, lane);
                    const int blk = lane % BPW;
                    if (blk < nvalid) part[PS * (w0 + blk) + lane / BPW] = sv;
                }
            } else if (lane < BPW && lane < nvalid) {
                double b0 = 0.0, b1 = 0.0;
                if (lane == 0)
                    for (int w = 0; w < NT / 64; w++) {
                        b0 += wl[16 + w];
                        b1 += wl[32 + w];
                    }
                part[PS * (w0 + lane)] = b0;
                part[PS * (w0 + lane) + 1] = b1;
            }
            if (lane < BPW && lane < nvalid) {
                double m = 0.0;
                for (int w = 0; w < NT / 64; w++) m = fmax(m, wl[w]);
                part[PS * (w0 + lane) + 2] = m;
            }
        }
        SY_PH_MARK(pht, ph);
        grid_barrier(A, bs, (nl + BPW - 1) / BPW);
        after_barrier();
        SY_PH_MARK(pht, ph + 1);
        // totals (wave 0).  Ordered sums: the partials staged in LDS by lane in chunks of 64 and
        // chained in block order in lanes 0 / 1.  Two-level variant (order-free): each lane adds
        // its blocks' partials, then a fixed shuffle tree.  The max is reduced across the wave.
        // Every workgroup forms the same totals in the same order.
        if (wv == 0 && PC) {
            double t0 = 0.0, t1 = 0.0, tm = 0.0;
            for (int b = lane; b < nl; b += 64) {
                t0 += load_published(part + PS * b);
                if (nsum > 1) t1 += load_published(part + PS * b + 1);
                tm = fmax(tm, load_published(part + PS * b + 2));
            }
            t0 = wave_sum(t0);
            t1 = wave_sum(t1);
            tm = wave_max(tm);
            if (lane == 0) {
                wl[48] = t0;
                wl[49] = t1;
                wl[50] = tm;
            }
        } else if (wv == 0) {
            double acc = 0.0, tm = 0.0;
            for (int base = 0; base < nl; base += 64) {
                const int cnt = nl - base < 64 ? nl - base : 64;
                const int cnt16 = (cnt + 15) & ~15;   // zero padding: acc starts at +0 and stays off -0
                if (lane < cnt16) {
                    const bool ok = lane < cnt;
                    red[lane] = ok ? load_published(part + PS * (base + lane)) : 0.0;
                    red[64 + lane] = ok && nsum > 1 ? load_published(part + PS * (base + lane) + 1) : 0.0;
                    if (ok) tm = fmax(tm, load_published(part + PS * (base + lane) + 2));
                }
                wave_lds_barrier();
                if (lane < (nsum > 1 ? 2 : 1)) acc = chain_n<200>(red + lane * 64, cnt16, acc);
                wave_lds_barrier();
            }
            tm = wave_max(tm);
            if (lane < 2) wl[48 + lane] = acc;
            if (lane == 0) wl[50] = tm;
        }
        __syncthreads();
        s0 = wl[48];
        s1 = wl[49];
        mx = wl[50];
        __syncthreads();
        SY_PH_MARK(pht, ph + 2);
    }
};

template <int D, int NB, bool PC>
__global__ void __launch_bounds__((BotCfg<D, NB>::NT)) k_bottom(BottomArgs A) {
    using G = Geo<D, NB>;
    using C = BotCfg<D, NB>;
    using B = Bottom<D, NB, PC>;
    constexpr int BPW = C::BPW, NT = C::NT, CPT = C::CPT, GPT = C::GPT;
    constexpr int PS = B::PS;
    if (A.skip && *A.skip) return;                 // every workgroup reads the same flag
    if (A.confine && (blockIdx.x & 7u)) return;   // one XCD: not a worker
    __shared__ int s_broken, s_light;     // a grid barrier timed out; one-XCD barriers (grid_barrier)
    if (threadIdx.x == 0) s_broken = s_light = 0;   // visible after the first __syncthreads below
    BarState bs{A.episode0, A.confine ? (int)(blockIdx.x >> 3) : (int)blockIdx.x, &s_broken, &s_light};
#ifdef SAYAKA_BOTTOM_PHASES
    __shared__ unsigned long long pht[1 + PH_NUM];   // [0] last mark, [1 + k] phase k
    if ((int)threadIdx.x <= PH_NUM) pht[threadIdx.x] = threadIdx.x == 0 ? wall_clock64() : 0;
#else
    unsigned long long* pht = nullptr;
#endif
    __shared__ __attribute__((aligned(16))) double sh[PC ? C::NSHARED_FAST : C::NSHARED];   // operand tiles, then reduction staging
    __shared__ double wl[64];
    __shared__ BlockInfo bi[BPW];
    __shared__ double dvs[BPW];
    __shared__ int gbs[BPW];            // global block id of each block of the workgroup
    __shared__ int gsrc[C::WGHOST];     // ghost slot -> source cell (global index); -1 zero (Neumann
                                        // wall / none); <= -2: Dirichlet wall, -(2 + interior tile index)
    __shared__ int gtile[C::WGHOST];    // ghost slot -> tile position
    __shared__ double rh[C::WCELL];     // rhat (constant over the iterations)
    __shared__ double vfs[C::WCELL];    // volume fractions of the workgroup's cells
    __shared__ double wsh[C::NWSH];     // face weights of the workgroup's blocks (when WLDS)
    // two-level preconditioner state (PC): coarse inverse, per-block sums of D r, D v, D p,
    // coarse values, per-wave block partials
    constexpr int NPC = PC ? PC_MAXBLK : 1;
    __shared__ float s0[PC ? PC_MAXBLK * (PC_MAXBLK + 1) : 1];
    __shared__ double sr_[NPC], sv_[NPC], sp_[NPC], ss_[NPC], cx_[NPC];
    __shared__ double wq[PC ? BPW * C::NW : 1];
    const StencilArgs& a = A.st;
    const int w0 = bs.wid * BPW;
    const int nvalid = a.nlist - w0 < BPW ? a.nlist - w0 : BPW;
    const int nl = a.nlist;
    const int ex = PC ? 0 : A.exact;
    if ((int)threadIdx.x < BPW) {
        const int q = threadIdx.x;
        gbs[q] = -1;
        if (q < nvalid) {
            const int b = a.list[w0 + q];
            gbs[q] = b;
            bi[q] = a.binfo[b];
            dvs[q] = a.met[b].dv;
        }
    }
    if constexpr (PC)
        for (int e = threadIdx.x; e < nl * nl; e += NT) s0[(e / nl) * (PC_MAXBLK + 1) + e % nl] = A.s0inv[e];
    __syncthreads();   // gbs / bi / dvs visible to every thread before they are used below
    if (C::WLDS)
        for (int e = threadIdx.x; e < C::NWSH; e += NT) {
            const int blk = e / (D * G::NFACE), rem = e % (D * G::NFACE);
            double wv = 0.0;
            if (blk < nvalid) {
                wv = a.w[gbs[blk] * (D * G::NFACE) + rem];
                // zero on Neumann wall faces (their face is on the block boundary)
                const int d = rem / G::NFACE, fi = rem % G::NFACE;
                const int fx = G::fx(d), fy = G::fy(d);
                const int cq[3] = {fi % fx, (fi / fx) % fy, fi / (fx * fy)};
                const int ext = d == 2 ? G::NBZ : NB;
                const int f = cq[d] == 0 ? 2 * d : (cq[d] == ext ? 2 * d + 1 : -1);
                if (f >= 0 && bi[blk].nbr[f] <= PHYS_BNDRY && bi[blk].bct[f] == BC_NEUMANN) wv = 0.0;
            }
            wsh[e] = wv;
        }
    for (int e = threadIdx.x; e < C::WCELL; e += NT) {
        const int blk = e / G::NCELL;
        vfs[e] = blk < nvalid ? a.vf[gbs[blk] * G::NCELL + e % G::NCELL] : 0.0;
    }
    __syncthreads();
    double* partA = A.partials;
    double* partB = A.partials + PS * a.nlist;
    double* partC = A.partials + 2 * PS * a.nlist;
    // own cells: e = tid + q*NT -> block e / NCELL of this workgroup, cell e % NCELL
    auto blk_of = [&](int q) { return (int)(threadIdx.x + q * NT) / G::NCELL; };
    auto cell_of = [&](int q) { return (int)(threadIdx.x + q * NT) % G::NCELL; };
    auto gidx = [&](int q) { return gbs[blk_of(q)] * G::NCELL + cell_of(q); };

    double x[CPT], r[CPT], p[CPT], v[CPT], s[CPT], t[CPT];
    double pG[GPT], vG[GPT], rG[GPT];   // ghost slot g = tid + u*NT
    double jo[CPT], jG[GPT];   // 1 / diag(A) on own / ghost cells (PC)
    int gpos[GPT];             // ghost source block's list position, -1 none (PC)
    double bq[BPW];                               // per-block partial sums of D * (vector) (PC)
    auto zero_bq = [&]() {
#pragma unroll
        for (int b = 0; b < BPW; b++) bq[b] = 0.0;
    };
    // thread's share of the per-block sums of D * y over its own cells
    auto block_terms = [&](const double (&y)[CPT]) {
        zero_bq();
#pragma unroll
        for (int q = 0; q < CPT; q++) {
            const int blk = blk_of(q);
            const double dy = blk < nvalid ? dvs[blk] * vfs[threadIdx.x + q * NT] * y[q] : 0.0;
#pragma unroll
            for (int b = 0; b < BPW; b++) bq[b] += blk == b ? dy : 0.0;
        }
    };

    double mxl = 0.0, a0 = 0.0, a1 = 0.0;
    int has_dir = 0;
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        x[q] = r[q] = p[q] = v[q] = s[q] = t[q] = 0.0;
        const int blk = blk_of(q);
        if constexpr (PC) jo[q] = 0.0;
        if (blk < nvalid) {
            const int gi = gidx(q);
            A.sol[gi] = 0.0;
            // r = rhs - A*x0 with x0 = 0 (mg_residual_level: 0 on solid cells)
            r[q] = a.vf[gi] > 0 ? A.rhs[gi] - 0.0 : 0.0;
            mxl = fmax(mxl, fabs(r[q]));
            if constexpr (PC) jo[q] = A.jac[(w0 + blk) * G::NCELL + cell_of(q)];
        }
        rh[threadIdx.x + q * NT] = r[q];
        B::stage(sh, ex, 0, blk, cell_of(q), r[q] * r[q], a0);
    }
#pragma unroll
    for (int u = 0; u < GPT; u++) {
        const int g = threadIdx.x + u * NT;
        pG[u] = vG[u] = rG[u] = 0.0;
        if constexpr (PC) {
            jG[u] = 0.0;
            gpos[u] = -1;
        }
        if (g >= C::WGHOST) continue;
        const int blk = g / G::NGHOST, gg = g % G::NGHOST;
        const int f = gg / G::NSLAB, pp = gg % G::NSLAB;
        const int d = f >> 1, sd = f & 1;
        int a1_, a2_;
        tang_axes<D>(d, a1_, a2_);
        const int ext = d == 2 ? G::NBZ : NB;
        int qc[3] = {0, 0, 0};
        qc[a1_] = pp % NB;
        if (D == 3) qc[a2_] = pp / NB;
        int tq[3] = {qc[0], qc[1], qc[2]};
        tq[d] = sd ? ext : -1;
        gtile[g] = blk * G::NTILE + G::tidx(tq[0], tq[1], tq[2]);
        int src = -1;
        if (blk < nvalid) {
            const int n = bi[blk].nbr[f];   // level-1 faces are same-level or walls
            if (n >= 0) {
                qc[d] = sd ? 0 : ext - 1;
                src = n * G::NCELL + G::cidx(qc[0], qc[1], qc[2]);
                const bool fluid = a.vf[src] > 0;
                rG[u] = fluid ? A.rhs[src] - 0.0 : 0.0;
                if constexpr (PC) {
                    const int pos = A.l1pos[n];
                    jG[u] = A.jac[pos * G::NCELL + (src - n * G::NCELL)];
                    gpos[u] = fluid ? pos : -1;
                }
            } else if (C::WLDS && n <= PHYS_BNDRY && bi[blk].bct[f] != BC_NEUMANN) {
                int iq[3] = {qc[0], qc[1], qc[2]};
                iq[d] = sd ? ext - 1 : 0;
                src = -2 - (blk * G::NTILE + G::tidx(iq[0], iq[1], iq[2]));
                has_dir = 1;
            }
        }
        gsrc[g] = src;
    }
    const int any_dir = __syncthreads_or(has_dir);
    double s0v, s1, mx;
    if constexpr (PC) {
        block_terms(r);
        B::publish_block_sums(bq, partC, wq, w0, nvalid);
    }
    B::reduce(pht, PH_INIT, bs, A, a0, a1, 1, mxl, partC, sh, wl, nvalid, s0v, s1, mx);
    if constexpr (PC) B::read_block_sums(partC, sr_, nl);
    const double rnorm0 = mx;
    double rho = s0v, rho1 = 0.0, alpha = 0.0, omega = 0.0;
    int iter = 0, conv = 0, stat = 0;
    // PC (the mixed cycle's coarse solve, a preconditioner): the residual after each full
    // iteration, the best one, and where it last halved.  BiCGSTAB can stagnate on a rare
    // right-hand side; the solve then stops once BOTTOM_STALL iterations have not halved the
    // best residual and returns its last iterate when that improved on the zero guess (the
    // FP64 modes keep the reference's rule: solution written on convergence only).
    double last_r = rnorm0, best_r = rnorm0;
    int best_it = 0, stall = 0;

    // operand tile of the workgroup's blocks from own registers + ghost registers
    auto fill_tile = [&](const double (&own)[CPT], const double (&gh)[GPT]) {
#pragma unroll
        for (int q = 0; q < CPT; q++) {
            int i, j, k;
            G::cell_ijk(cell_of(q), i, j, k);
            sh[blk_of(q) * G::NTILE + G::tidx(i, j, k)] = own[q];
        }
#pragma unroll
        for (int u = 0; u < GPT; u++) {
            const int g = threadIdx.x + u * NT;
            if (g < C::WGHOST && gsrc[g] >= -1) sh[gtile[g]] = gsrc[g] >= 0 ? gh[u] : 0.0;
        }
        __syncthreads();
        if (any_dir) {   // Dirichlet / SimpleFill wall ghosts = -phi_c of the adjacent cell
#pragma unroll
            for (int u = 0; u < GPT; u++) {
                const int g = threadIdx.x + u * NT;
                if (g < C::WGHOST && gsrc[g] <= -2) sh[gtile[g]] = -sh[-2 - gsrc[g]];
            }
            __syncthreads();
        }
    };
    auto apply = [&](double (&out)[CPT]) {
#pragma unroll
        for (int q = 0; q < CPT; q++) {
            out[q] = 0.0;
            const int blk = blk_of(q);
            if (blk >= nvalid) continue;
            int i, j, k;
            G::cell_ijk(cell_of(q), i, j, k);
            if (C::WLDS)
                out[q] = bot_apply_cell<D, NB>(sh + blk * G::NTILE, wsh + blk * (D * G::NFACE), i, j, k, a.alpha,
                                               dvs[blk], vfs[threadIdx.x + q * NT]);
            else {
                double wr[6] = {0, 0, 0, 0, 0, 0}, wcr[6] = {0, 0, 0, 0, 0, 0};
                const double* wb = a.w + gbs[blk] * (D * G::NFACE);
#pragma unroll
                for (int f = 0; f < G::NF; f++)
                    wr[f] = wb[(f >> 1) * G::NFACE + G::fidx(f >> 1, i + (f == 1), j + (f == 3), k + (f == 5))];
                out[q] = cell_op<D, NB, false, double>(sh + blk * G::NTILE, wr, wcr, i, j, k, bi[blk], 1, a.alpha,
                                                       dvs[blk], vfs[threadIdx.x + q * NT], 0.0);
            }
        }
        __syncthreads();   // tile free for the reduction staging
    };
    auto load_ghosts = [&](double (&dst)[GPT], const double* src) {
#pragma unroll
        for (int u = 0; u < GPT; u++) {
            const int g = threadIdx.x + u * NT;
            if (g < C::WGHOST && gsrc[g] >= 0) dst[u] = load_published(src + gsrc[g]);   // another worker's
        }
    };
    // preconditioned values M^-1 y on own cells (into out) and ghost cells (into outG) from
    // the coarse values cx_ of every block (PC)
    auto precondition = [&](const double (&y)[CPT], const double (&yG)[GPT], double (&out)[CPT],
                            double (&outG)[GPT]) {
#pragma unroll
        for (int q = 0; q < CPT; q++) {
            const int blk = blk_of(q);
            out[q] = 0.0;
            if (blk < nvalid && vfs[threadIdx.x + q * NT] > 0) out[q] = jo[q] * y[q] + cx_[w0 + blk];
        }
#pragma unroll
        for (int u = 0; u < GPT; u++) outG[u] = gpos[u] >= 0 ? jG[u] * yG[u] + cx_[gpos[u]] : 0.0;
    };

    if (rnorm0 != 0.0) {
        for (iter = 1; iter <= A.max_iter; iter++) {
            if (rho == 0.0) { stat = 1; break; }
            const double beta = iter == 1 ? 0.0 : (rho / rho1) * (alpha / omega);
#pragma unroll
            for (int q = 0; q < CPT; q++) {
                if (iter == 1) p[q] = r[q];
                else {
                    const double p1 = 1.0 * p[q] + -omega * v[q];
                    p[q] = 1.0 * r[q] + beta * p1;
                }
            }
#pragma unroll
            for (int u = 0; u < GPT; u++) {
                if (iter == 1) pG[u] = rG[u];
                else {
                    const double p1 = 1.0 * pG[u] + -omega * vG[u];
                    pG[u] = 1.0 * rG[u] + beta * p1;
                }
            }
            // v = A phat ; alpha_den = rhat . v  (barrier A, publishes v)
            if constexpr (PC) {
                // P^T D p = P^T D r + beta (P^T D p_old - omega P^T D v); phat = M^-1 p (into t)
                for (int b = threadIdx.x; b < nl; b += NT)
                    sp_[b] = iter == 1 ? sr_[b] : sr_[b] + beta * (sp_[b] - omega * sv_[b]);
                __syncthreads();
                B::coarse_mv(s0, sp_, cx_, nl);
                double phG[GPT];
                precondition(p, pG, t, phG);
                fill_tile(t, phG);
            } else {
                fill_tile(p, pG);
            }
            apply(v);
            a0 = 0.0;
#pragma unroll
            for (int q = 0; q < CPT; q++) {
                const int blk = blk_of(q);
                B::stage(sh, ex, 0, blk, cell_of(q), rh[threadIdx.x + q * NT] * v[q], a0);
                if (blk < nvalid) A.vpub[gidx(q)] = v[q];
            }
            if constexpr (PC) {
                block_terms(v);
                B::publish_block_sums(bq, partA, wq, w0, nvalid);
            }
            SY_PH_MARK(pht, PH_OP_V);
            B::reduce(pht, PH_A, bs, A, a0, 0.0, 1, 0.0, partA, sh, wl, nvalid, s0v, s1, mx,
                      [&]() { load_ghosts(vG, A.vpub); });
            if constexpr (PC) B::read_block_sums(partA, sv_, nl);
            if (s0v == 0.0) { stat = 1; break; }
            alpha = rho / s0v;
            // s = r - alpha v ; x += alpha phat ; t = A shat ; |s|, t.s, t.t  (barrier B)
            mxl = 0.0;
#pragma unroll
            for (int q = 0; q < CPT; q++) {
                s[q] = 1.0 * r[q] + -alpha * v[q];
                x[q] = 1.0 * x[q] + alpha * (PC ? t[q] : p[q]);
                if (blk_of(q) < nvalid) mxl = fmax(mxl, fabs(s[q]));
            }
            double shat[CPT];
            {
                double sG[GPT];
#pragma unroll
                for (int u = 0; u < GPT; u++) sG[u] = 1.0 * rG[u] + -alpha * vG[u];
                if constexpr (PC) {
                    for (int b = threadIdx.x; b < nl; b += NT) ss_[b] = sr_[b] - alpha * sv_[b];   // P^T D s
                    __syncthreads();
                    B::coarse_mv(s0, ss_, cx_, nl);
                    double shG[GPT];
                    precondition(s, sG, shat, shG);
                    fill_tile(shat, shG);
                } else {
                    fill_tile(s, sG);
                }
            }
            apply(t);
            a0 = a1 = 0.0;
#pragma unroll
            for (int q = 0; q < CPT; q++) {
                B::stage(sh, ex, 0, blk_of(q), cell_of(q), t[q] * s[q], a0);
                B::stage(sh, ex, 1, blk_of(q), cell_of(q), t[q] * t[q], a1);
            }
            double ts, tt;
            SY_PH_MARK(pht, PH_OP_T);
            B::reduce(pht, PH_B, bs, A, a0, a1, 2, mxl, partB, sh, wl, nvalid, ts, tt, mx);
            if (mx < rnorm0 * A.tol_rel || mx < A.tol_abs) { conv = 1; break; }
            if (tt == 0.0) { stat = 1; break; }
            omega = ts / tt;
            // x += omega shat ; r = s - omega t ; |r|, rhat . r  (barrier C, publishes r)
            mxl = 0.0;
            a0 = 0.0;
#pragma unroll
            for (int q = 0; q < CPT; q++) {
                x[q] = 1.0 * x[q] + omega * (PC ? shat[q] : s[q]);
                r[q] = 1.0 * s[q] + -omega * t[q];
                const int blk = blk_of(q);
                B::stage(sh, ex, 0, blk, cell_of(q), rh[threadIdx.x + q * NT] * r[q], a0);
                if (blk < nvalid) {
                    mxl = fmax(mxl, fabs(r[q]));
                    A.rpub[gidx(q)] = r[q];
                }
            }
            if constexpr (PC) {
                block_terms(r);
                B::publish_block_sums(bq, partC, wq, w0, nvalid);
            }
            SY_PH_MARK(pht, PH_UPD_R);
            B::reduce(pht, PH_C, bs, A, a0, 0.0, 1, mxl, partC, sh, wl, nvalid, s0v, s1, mx,
                      [&]() { load_ghosts(rG, A.rpub); });
            if constexpr (PC) B::read_block_sums(partC, sr_, nl);
            SY_PH_MARK(pht, PH_R_GHOST);
            if (mx < rnorm0 * A.tol_rel + A.tol_abs) { conv = 1; break; }
            if constexpr (PC) {
                last_r = mx;
                if (mx < 0.5 * best_r) {
                    best_r = mx;
                    best_it = iter;
                } else if (iter - best_it >= BOTTOM_STALL) {
                    stall = 1;
                    break;
                }
            }
            if (omega == 0.0) { stat = 1; break; }
            rho1 = rho;
            rho = s0v;
        }
        if (iter > A.max_iter) iter = A.max_iter;
    }
    if (PC && !conv && !stat && last_r < rnorm0) conv = 1;   // stalled or out of iterations: last iterate
    if (conv) {
#pragma unroll
        for (int q = 0; q < CPT; q++)
            if (blk_of(q) < nvalid) A.sol[gidx(q)] = 0.0 + x[q];
    }
#ifdef SAYAKA_BOTTOM_PHASES
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        for (int k = 0; k < PH_NUM; k++) g_bottom_phases[k] += pht[1 + k];
        g_bottom_phases[30] += iter;
        g_bottom_phases[31] += 1;
    }
#endif
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        A.stats[5] += A.confine;
        A.stats[6] += A.confine && !s_light;
        A.stats[0] += 1;
        A.stats[1] += iter;
        A.stats[2] = iter;
        A.stats[3] = stat;
        A.stats[7] += stall;
    }
}

}  // namespace sydev

namespace sydev {

// Data of the two-level preconditioner for the level-1 problem, per level-1 block (one
// workgroup each, list position = blockIdx.x): 1 / diag(A) per cell (dv * vf over the sum
// of the cell's face weights, with the operator's wall treatment: Neumann faces 0,
// Dirichlet / SimpleFill faces doubled), and per block face the sum of the weights crossing
// it (an inter-block coupling of S0, or on a wall the wall's share of S0's diagonal).
template <int D, int NB>
__global__ void __launch_bounds__((Geo<D, NB>::THREADS)) k_l1_coarse(StencilArgs a, double* jac, double* fsum) {
    using G = Geo<D, NB>;
    constexpr int NT = G::THREADS, NW = (NT + 63) / 64;
    __shared__ double red[6][NW];
    const int pos = blockIdx.x;
    const int b = a.list[pos];
    const double dv = a.met[b].dv;
    const double* wb = a.w + b * (D * G::NFACE);
    const BlockInfo& bi = a.binfo[b];
    int nbr[6], bct[6];
#pragma unroll
    for (int f = 0; f < 6; f++) {
        nbr[f] = f < G::NF ? bi.nbr[f] : 0;
        bct[f] = f < G::NF ? bi.bct[f] : -1;
    }
    double fs[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int c = threadIdx.x; c < G::NCELL; c += NT) {
        int i, j, k;
        G::cell_ijk(c, i, j, k);
        double sw = 0.0;
#pragma unroll
        for (int f = 0; f < G::NF; f++) {
            const int d = f >> 1;
            double w = wb[d * G::NFACE + G::fidx(d, i + (f == 1), j + (f == 3), k + (f == 5))];
            const int ext = d == 2 ? G::NBZ : NB;
            const int cc = d == 0 ? i : (d == 1 ? j : k);
            const bool bnd = (f & 1) ? cc == ext - 1 : cc == 0;
            if (bnd && nbr[f] <= PHYS_BNDRY) w = bct[f] == BC_NEUMANN ? 0.0 : 2.0 * w;
            sw += w;
            if (bnd) fs[f] += w;
        }
        const double vf = a.vf[b * G::NCELL + c];
        jac[pos * G::NCELL + c] = (vf > 0 && sw > 0) ? dv * vf / sw : 0.0;
    }
#pragma unroll
    for (int f = 0; f < 6; f++) {
        const double v = wave_sum(fs[f]);
        if ((threadIdx.x & 63) == 0) red[f][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        double t = 0.0;
        for (int w = 0; w < NW; w++) t += red[threadIdx.x][w];
        fsum[pos * 6 + threadIdx.x] = t;
    }
}

}  // namespace sydev
