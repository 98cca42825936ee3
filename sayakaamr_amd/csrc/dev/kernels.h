// kernels.h -- HIP kernels of the MAC solve, templated on dimension D and block edge NB.
//
// Arithmetic is written to reproduce the reference operation order (files cited per
// kernel) so that, built with -ffp-contract=off, the FP64 kernels are bit-identical to
// the CPU oracle; only reductions (dot products) sum in a different, fixed order.
#pragma once

#include "chain_asm.h"
#include "dev_common.h"

namespace sydev {

// ------------------------------------------------------------------ reductions (fixed order)
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Block-wide reduction of up to three values (two sums, one max), deterministic for a fixed
// block size.  Result valid in thread 0.
template <int NT>
__device__ inline void block_reduce3(double& s0, double& s1, double& mx, double* lds) {
    constexpr int NW = (NT + 63) / 64;
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    mx = wave_max(mx);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        lds[wv] = s0;
        lds[NW + wv] = s1;
        lds[2 * NW + wv] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0, m = 0.0;
        for (int w = 0; w < NW; w++) {
            a += lds[w];
            b += lds[NW + w];
            m = fmax(m, lds[2 * NW + w]);
        }
        s0 = a;
        s1 = b;
        mx = m;
    }
    __syncthreads();
}

// Scalars of the Krylov loops kept on the device.
enum Scalar {
    S_RHO = 0, S_RHO1, S_ALPHA, S_OMEGA, S_ALPHA_DEN, S_TS, S_TT, S_SNORM, S_RNORM, S_RNORM0,
    S_DOT, S_NORM, S_AUX0, S_AUX1, S_NSCALAR
};

// What the last workgroup computes from the global sums.
enum Finalize {
    FIN_NONE = 0,
    FIN_INIT,       // rnorm0 = max, rho = sum0
    FIN_ALPHA,      // alpha_den = sum0, alpha = rho / alpha_den
    FIN_SNORM,      // snorm = max
    FIN_OMEGA,      // ts = sum0, tt = sum1, omega = ts / tt
    FIN_RUPDATE,    // rnorm = max, rho1 = rho, rho = sum0
    FIN_DOT,        // dot = sum0
    FIN_NORM,       // norm = max
    FIN_DOTNORM     // dot = sum0, aux0 = sum1, norm = max
};

// Grid reductions run in two launches: every workgroup of the operator / vector kernel
// stores its (sum, sum, max) partials with plain stores, and a one-workgroup kernel
// (k_red_finish) sums them (in list order when exact) and updates the Krylov scalars.  The
// single-launch alternative -- each workgroup publishing behind a device-scope fence and an
// atomic counter electing the last one -- measured ~35 ns per workgroup on MI355X (a
// device-scope release makes the XCD's L2 write its dirty lines back so that the other XCDs
// see them): 0.15 ms on the 4 264 blocks of a 2 M-cell tower, 1.9 ms on 38 848 blocks,
// against a few microseconds for the extra launch.
struct RedBuf {
    double* partials;   // 3 per workgroup
    double* scal;       // S_NSCALAR doubles
    int exact;          // 1: sums in the reference's order (cells k,j,i in a block, then
                        //    blocks in list order: DotProd / mg_dotprod_level,
                        //    SayakaBlockData.cpp:185-208, SayakaMacUtil.cpp:79-106)
    // Stopping tests of a solve's host, evaluated on the device (the same expressions on the
    // same scalars) for the launches it queues before reading them: stop[0] (FIN_SNORM) = 1 when
    // the half step stops the solve (alpha_den == 0, |s| < rnorm0 tol_rel, |s| < tol_abs);
    // stop[1] (FIN_RUPDATE) = 1 when the half step or the full step does (t.t == 0,
    // |r| < rnorm0 tol_rel + tol_abs, omega == 0, or the next iteration's rho == 0).
    int* stop = nullptr;
    double tol_rel = 0.0;
    double tol_abs = 0.0;
    // FIN_RUPDATE: a copy of the scalars in host memory (device address of coherent pinned
    // memory), then seq in its slot S_NSCALAR, for a host that does not wait for the stream
    double* host = nullptr;
    unsigned long long seq = 0;
};

// LDS byte address of a pointer into __shared__ memory.
__device__ inline unsigned lds_addr(const double* p) {
    return (unsigned)(uintptr_t)(const __attribute__((address_space(3))) double*)p;
}

// Sequential sum of n (a multiple of 16) contiguous 16-byte aligned LDS values onto s, in index
// order, as the assembly blocks of chain_asm.h (blocks of 512, 64 and 16 values; BASE = the
// first of the 36 registers a block uses, 16 or 200).  A sequential FP64 sum is bound by its
// dependent adds and by the LDS reads that feed them: 10.9 clocks per add on MI355X against 16.3
// for the compiler's own code, which kept two or three reads ahead in register-heavy kernels
// (3.7 us per 512 values inside k_bottom) and none in one build (7.6 us).
template <int BASE = 16>
__device__ inline double lds_chain(const double* p, int n, double s) {
    const unsigned a = lds_addr(p);
    int k = 0;
    for (; k + 512 <= n; k += 512) s = lds_chain_asm<512, BASE>(a + 8u * k, s);
    for (; k + 64 <= n; k += 64) s = lds_chain_asm<64, BASE>(a + 8u * k, s);
    for (; k + 16 <= n; k += 16) s = lds_chain_asm<16, BASE>(a + 8u * k, s);
    return s;
}

// Lane l of the calling wave adds the N values staged for chain l onto s, in index order.  Chain
// l occupies red[l*(N+2) .. l*(N+2)+N): the stride keeps every chain 16-byte aligned and puts
// concurrent chains on distinct LDS banks.
template <int N, int BASE = 16>
__device__ inline double lane_chain(const double* red, int l, double s = 0.0) {
    static_assert(N % 16 == 0, "chain length must be a multiple of 16");
    return lds_chain<BASE>(red + l * (N + 2), N, s);
}

// Sequential sum of n (a multiple of 16) contiguous 16-byte aligned LDS values onto acc.
template <int BASE = 16>
__device__ inline double chain_n(const double* p, int n, double acc) {
    if (n <= 0) return acc;
    return lds_chain<BASE>(p, n, acc);
}

// LDS written by lanes of this wave becomes visible to the wave's other lanes (no workgroup
// barrier: the other waves are not taking part).
__device__ inline void wave_lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline void stage_chain(double* red, int n, int l, int c, double v) { red[l * (n + 2) + c] = v; }

// Sums, in cell order, of nsum (1 or 2) per-cell values held by the threads (cell
// c = tid + q*NT); the two chains run concurrently in lanes 0 and 1.  The cells of one q are
// the contiguous run q*NT .. q*NT+NT-1, so the chains go through LDS one run at a time, the
// sums carried from run to run: `scratch` is 16-byte aligned LDS of at least 2*min(NCELL,
// NT)+4 doubles.  Results valid in thread 0.
template <int NT, int NCELL, int CPT>
__device__ inline void seq_block_sums(const double (&v0)[CPT], const double (&v1)[CPT], int nsum, double* scratch,
                                      double& s0, double& s1) {
    static_assert(CPT == 1 || NCELL == CPT * NT, "runs of NT cells");
    constexpr int RUN = CPT == 1 ? NCELL : NT;
    double r = 0.0;
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        __syncthreads();
        if ((int)threadIdx.x < RUN) {
            stage_chain(scratch, RUN, 0, threadIdx.x, v0[q]);
            if (nsum > 1) stage_chain(scratch, RUN, 1, threadIdx.x, v1[q]);
        }
        __syncthreads();
        if ((int)threadIdx.x < nsum) r = lane_chain<RUN>(scratch, threadIdx.x, r);
    }
    __syncthreads();
    if (threadIdx.x == 1) scratch[0] = r;
    __syncthreads();
    s0 = r;
    s1 = nsum > 1 ? scratch[0] : 0.0;
    __syncthreads();
}

// Coherent load of a value another workgroup published before a device-scope fence.
__device__ inline double load_published(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Staged layout of the two 512-term sums of seg_chain512_asm: term c of sum h at
// h * SEG_STRIDE + c + 2 (c / 16), so that the lanes' 128-byte runs start 144 bytes apart and
// spread over the LDS banks.
constexpr int SEG_STRIDE = 512 + 512 / 8 + 2;
__device__ inline int seg_slot(int h, int c) { return h * SEG_STRIDE + c + 2 * (c >> 4); }

// Sum of n (sum, sum, max) partials written by an earlier launch: exact mode sums in index
// order, fast mode by a strided tree.  The max is order-free.  Results valid in thread 0.
// Exact mode (NT = 512): chunks of 512 partials, wave 0 chaining chunk k (both sums at once,
// seg_chain512_asm, the running sums carried from chunk to chunk) while the other waves load
// chunk k + 1 into the other of two LDS buffers; zero padding after the last partial (the
// running sums start at +0 and never become -0, so adding +0 changes nothing).
template <int NT>
__device__ inline void sum_published(const double* part, unsigned n, bool exact, int nsum, double& a, double& b,
                                     double& m, double* chunk /* 4 * SEG_STRIDE doubles of LDS, 16-byte aligned */,
                                     double* lds) {
    static_assert(NT == 512, "one partial per thread and chunk");
    a = 0.0;
    b = 0.0;
    m = 0.0;
    if (exact) {
        double ml = 0.0;
        const unsigned nchunk = (n + 511u) / 512u;
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        // stage chunk k (partials 512 k ..) into buffer k % 2 with threads t0 .. t0 + nt - 1
        auto stage = [&](unsigned k, int t, int nt) {
            double* buf = chunk + (k & 1u) * 2 * SEG_STRIDE;
            for (int g = t; g < 512; g += nt) {
                const unsigned i = 512u * k + (unsigned)g;
                const bool ok = i < n;
                buf[seg_slot(0, g)] = ok ? part[3 * i] : 0.0;
                buf[seg_slot(1, g)] = ok && nsum > 1 ? part[3 * i + 1] : 0.0;
                if (ok) ml = fmax(ml, part[3 * i + 2]);
            }
        };
        if (nchunk > 0) stage(0, threadIdx.x, NT);
        __syncthreads();
        double acc = 0.0;   // wave 0: sum 0 in lanes 0-31, sum 1 in lanes 32-63 (carried in lanes 31, 63)
        for (unsigned k = 0; k < nchunk; k++) {
            if (wv == 0) {
                const double* buf = chunk + (k & 1u) * 2 * SEG_STRIDE;
                const double r = seg_chain512_asm<16>(lds_addr(buf + seg_slot(lane >> 5, 16 * (lane & 31))), acc);
                acc = __shfl(r, lane < 32 ? 31 : 63, 64);
            } else if (k + 1 < nchunk) {
                stage(k + 1, threadIdx.x - 64, NT - 64);
            }
            __syncthreads();
        }
        if (threadIdx.x == 32) chunk[0] = acc;   // sum 1 (the buffers are no longer read)
        __syncthreads();
        const double sb = chunk[0];
        double z0 = 0.0, z1 = 0.0;
        block_reduce3<NT>(z0, z1, ml, lds);   // contains barriers
        a = acc;
        b = nsum > 1 ? sb : 0.0;
        m = ml;
    } else {
        for (unsigned g = threadIdx.x; g < n; g += NT) {
            a += part[3 * g];
            b += part[3 * g + 1];
            m = fmax(m, part[3 * g + 2]);
        }
        block_reduce3<NT>(a, b, m, lds);
    }
}

// Per-workgroup totals (valid in thread 0) to the partials of this workgroup.
__device__ inline void grid_finish(double s0, double s1, double mx, const RedBuf& rb) {
    if (threadIdx.x == 0) {
        double* p = rb.partials + 3 * (int)blockIdx.x;
        p[0] = s0;
        p[1] = s1;
        p[2] = mx;
    }
}

// Second launch of a grid reduction (one workgroup of RED_NT threads): the n partials of the
// previous launch summed (in list order when exact) and the Krylov scalars updated.
constexpr int RED_NT = 512;
__global__ void __launch_bounds__(RED_NT) k_red_finish(RedBuf rb, int n, int nsum, int fin) {
    __shared__ __attribute__((aligned(16))) double chunk[4 * SEG_STRIDE];
    __shared__ double lds[3 * (RED_NT / 64)];
    double a, b, m;
    sum_published<RED_NT>(rb.partials, (unsigned)n, rb.exact != 0, nsum, a, b, m, chunk, lds);
    if (threadIdx.x == 0) {
        double* S = rb.scal;
        switch (fin) {
            case FIN_INIT: S[S_RNORM0] = m; S[S_RHO] = a; break;
            case FIN_ALPHA: S[S_ALPHA_DEN] = a; S[S_ALPHA] = a != 0.0 ? S[S_RHO] / a : 0.0; break;
            case FIN_SNORM:
                S[S_SNORM] = m;
                if (rb.stop) rb.stop[0] = (S[S_ALPHA_DEN] == 0.0 || m < S[S_RNORM0] * rb.tol_rel || m < rb.tol_abs) ? 1 : 0;
                break;
            case FIN_OMEGA: S[S_TS] = a; S[S_TT] = b; S[S_OMEGA] = b != 0.0 ? a / b : 0.0; break;
            case FIN_RUPDATE:
                S[S_RNORM] = m;
                S[S_RHO1] = S[S_RHO];
                S[S_RHO] = a;
                if (rb.stop)
                    rb.stop[1] = (rb.stop[0] || S[S_TT] == 0.0 || m < S[S_RNORM0] * rb.tol_rel + rb.tol_abs ||
                                  S[S_OMEGA] == 0.0 || a == 0.0) ? 1 : 0;
                break;
            case FIN_DOT: S[S_DOT] = a; break;
            case FIN_NORM: S[S_NORM] = m; break;
            case FIN_DOTNORM: S[S_DOT] = a; S[S_AUX0] = b; S[S_NORM] = m; break;
            default: break;
        }
        if (rb.host) {   // the values first, then seq with a system-scope release
            double v[S_NSCALAR];   // every load issued before the first store
#pragma unroll
            for (int k = 0; k < S_NSCALAR; k++) v[k] = S[k];
#pragma unroll
            for (int k = 0; k < S_NSCALAR; k++) rb.host[k] = v[k];
            __hip_atomic_store(reinterpret_cast<unsigned long long*>(rb.host + S_NSCALAR), rb.seq, __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// Block totals of nsum per-cell sums (q0, q1) and a max, stored as the workgroup's
// partials.  `scratch` is 16-byte aligned LDS of at least 2*min(NCELL, NT)+4 doubles.
template <int NT, int NCELL, int CPT>
__device__ inline void grid_reduce_cells(const double (&q0)[CPT], const double (&q1)[CPT], double mx, int nsum,
                                         const RedBuf& rb, double* scratch, double* lds) {
    double s0 = 0.0, s1 = 0.0;
    if (rb.exact) {
        if (nsum > 0) seq_block_sums<NT, NCELL, CPT>(q0, q1, nsum, scratch, s0, s1);
        double z0 = 0.0, z1 = 0.0;
        block_reduce3<NT>(z0, z1, mx, lds);
    } else {
#pragma unroll
        for (int q = 0; q < CPT; q++) {
            s0 += q0[q];
            s1 += q1[q];
        }
        block_reduce3<NT>(s0, s1, mx, lds);
    }
    grid_finish(s0, s1, mx, rb);
}

// ---------------------------------------------------------------------------- stencil core
template <class T>
struct StencilArgsT {
    const int* list;          // block ids processed by this launch
    int nlist;
    int m;                    // MG tower level (face types depend on it)
    const BlockInfo* binfo;
    const BlockMetric* met;
    const T* w;               // own face weights [blk][D][NFACE]
    const T* wcf;             // coarse-side coarse/fine weights [blk][NF][NSLAB]
    const T* vf;              // [blk][NCELL]
    const T* x;               // input field (smoother: updated in place)
    T* out;                   // apply / residual output
    const T* rhs;             // residual / smoother right-hand side
    const T* dotv;            // second operand of the fused dot (may be null)
    T* out2;                  // second output (residual init: rhat)
    T* out3;                  // third output (residual init: x zeroed)
    int colour;
    T alpha;                  // MAC alpha (0 in every configuration, SayakaMacOperator.cpp:1268)
};
using StencilArgs = StencilArgsT<double>;

enum StencilMode {
    ST_APPLY = 0, ST_RESID = 1, ST_SMOOTH = 2, ST_APPLY_DOT = 3, ST_APPLY_DOT2 = 4, ST_RESID_INIT = 5,
    // one colour of a red-black sweep from x into out (cells of the other colour copied):
    // blocks of several sub-levels can share a launch, since a block reading a coarse/fine
    // neighbour sees that neighbour's values from before the launch, whatever the order in
    // which workgroups run (the mixed-precision cycle's smoother)
    ST_SMOOTH_DB = 6
};

template <int D, int NB>
__device__ inline int face_type(const BlockInfo& bi, int f, int m) {
    const int n = bi.nbr[f];
    if (n <= PHYS_BNDRY) return FT_BC;
    if (n < 0) return FT_FC;
    if (bi.level < m && !bi.nbr_leaf[f]) return FT_CF;
    return FT_FF;
}

// Sum of the fine-side contributions w_f * phi_f over the 2^(d-1) fine faces that cover
// one coarse face cell, in the order of restrictFaceBlockData_Average
// (SayakaInterpPatch.cpp:659-687): higher tangential axis outer.  These are the values
// the reference saves on FINE_CRSE faces (SayakaMacOperator.cpp:1446-1455, 1310-1319) and
// sums into the coarse block's CRSE_FINE face (SayakaTower.cpp:360-413).  The fine block
// comes from the table (BlockInfo::nfc), so no dependent table load is needed.
template <int D, int NB, class T>
__device__ inline T cf_sum(int f, int t1, int t2, const BlockInfo& bi, const T* __restrict__ x,
                           const T* __restrict__ w, const T* __restrict__ xslab = nullptr) {
    using G = Geo<D, NB>;
    const int d = f >> 1, s = f & 1;
    int a1, a2;
    tang_axes<D>(d, a1, a2);
    const int half = NB / 2;
    const int s1 = t1 >= half ? 1 : 0;
    const int s2 = (D == 3 && t2 >= half) ? 1 : 0;
    const int fb = bi.nfc[f][s1 + 2 * s2];
    const int ext = d == 2 ? G::NBZ : NB;
    T acc = T(0);
    for (int q2 = 0; q2 <= (D == 3 ? 1 : 0); q2++) {
        for (int q1 = 0; q1 <= 1; q1++) {
            int c[3] = {0, 0, 0}, fq[3] = {0, 0, 0};
            c[d] = s ? 0 : ext - 1;
            fq[d] = s ? 0 : ext;
            c[a1] = fq[a1] = (t1 - s1 * half) * 2 + q1;
            if (D == 3) c[a2] = fq[a2] = (t2 - s2 * half) * 2 + q2;
            const T wv = w[(fb * D + d) * G::NFACE + G::fidx(d, fq[0], fq[1], fq[2])];
            const T ph = (d == 0 && xslab) ? xslab[(fb * 2 + (s ? 0 : 1)) * G::NSLAB + c[1] + NB * c[2]]
                                           : x[fb * G::NCELL + G::cidx(c[0], c[1], c[2])];
            acc += wv * ph;
        }
    }
    return acc;
}

// The one-cell face halo of a block into LDS (the block's own cells are already there).
// Same-level neighbour: its boundary slab.  Coarser neighbour: the coarse leaf cell that the
// reference's injection from the parent's ghost resolves to (SayakaFillPatch.cpp:192-299,
// SayakaTreeData.h:41-44).  Finer neighbour on tower m (a coarse/fine face seen from the
// coarse side): the sum of the fine w*phi over the face cell (cf_sum), which is all the
// operator needs from that side.  Walls: 0 (walls enter through the weights).
// xslab (optional): every block's x-boundary columns of x, [block][side i = 0 / NB-1][j + NB k]
// (written by the kernel that wrote x), read for the x faces instead of x itself: a column of
// x is one cache line per cell, the slab one contiguous run.
template <int D, int NB, class T, int NTH = Geo<D, NB>::THREADS>
__device__ inline void load_halo(T* tile, const T* __restrict__ x, const T* __restrict__ w, const BlockInfo& bi,
                                 int m, const T* __restrict__ xslab = nullptr) {
    using G = Geo<D, NB>;
    // stride by the compile-time thread count: blockDim is read from the dispatch packet,
    // one scalar load per wave that measured ~14 us per launch on MI355X
    for (int g = threadIdx.x; g < G::NGHOST; g += NTH) {
        const int f = g / G::NSLAB, p = g % G::NSLAB;
        const int d = f >> 1, s = f & 1;
        int a1, a2;
        tang_axes<D>(d, a1, a2);
        const int ext = d == 2 ? G::NBZ : NB;
        int q[3] = {0, 0, 0};
        q[a1] = p % NB;
        if (D == 3) q[a2] = p / NB;
        int gq[3] = {q[0], q[1], q[2]};
        gq[d] = s ? ext : -1;
        const int n = bi.nbr[f];
        T v = T(0);
        const bool sl = d == 0 && xslab;
        if (n >= 0) {
            if (bi.level < m && !bi.nbr_leaf[f]) {
                v = cf_sum<D, NB, T>(f, q[a1], D == 3 ? q[a2] : 0, bi, x, w, xslab);
            } else {
                q[d] = s ? 0 : ext - 1;
                v = sl ? xslab[(n * 2 + (s ? 0 : 1)) * G::NSLAB + p] : x[n * G::NCELL + G::cidx(q[0], q[1], q[2])];
            }
        } else if (n > PHYS_BNDRY) {
            const int cb = bi.cnbr[f];
            const int wc = bi.which_child;
            int r[3] = {0, 0, 0};
            r[d] = s ? 0 : ext - 1;
            r[a1] = (q[a1] >> 1) + side_of(wc, a1) * (NB / 2);
            if (D == 3) r[a2] = (q[a2] >> 1) + side_of(wc, a2) * (NB / 2);
            v = sl ? xslab[(cb * 2 + (s ? 0 : 1)) * G::NSLAB + r[1] + NB * r[2]]
                   : x[cb * G::NCELL + G::cidx(r[0], r[1], r[2])];
        }
        tile[G::tidx(gq[0], gq[1], gq[2])] = v;
    }
}

// One cell of the operator (MODE apply) or of a Gauss-Seidel update (MODE smooth):
// mg_apply_block_usefill (SayakaMacOperator.cpp:1372-1458) and
// mg_smooth_block_rbcycle_usefill (SayakaMacOperator.cpp:1232-1320) with the face
// coefficients of cell_eval_coef (1115-1167) already folded into the weights: wr[f] is the
// own weight of face f of this cell, wcr[f] the coarse-side weight of a coarse/fine face, and
// the tile's halo holds the fine sum on such faces (load_halo).
template <int D, int NB, bool SMOOTH, class T>
__device__ inline T cell_op(const T* tile, const T (&wr)[6], const T (&wcr)[6], int i, int j, int k,
                            const BlockInfo& bi, int m, T alpha, T dv, T vfv, T rhsv) {
    using G = Geo<D, NB>;
    T diag = alpha * dv;
    T offdiag = T(0);
    const T phic = tile[G::tidx(i, j, k)];
#pragma unroll
    for (int f = 0; f < G::NF; f++) {
        const int d = f >> 1, s = f & 1;
        const int idx = d == 0 ? i : (d == 1 ? j : k);
        const int ext = d == 2 ? G::NBZ : NB;
        const bool bnd = s ? idx == ext - 1 : idx == 0;
        const int ni = G::tidx(i + (f == 1) - (f == 0), j + (f == 3) - (f == 2), k + (f == 5) - (f == 4));
        T off, offval;
        if (!bnd) {
            off = wr[f];
            offval = off * tile[ni];
        } else {
            const int ft = face_type<D, NB>(bi, f, m);
            if (ft == FT_CF) {
                off = wcr[f];
                offval = tile[ni];
            } else if (ft == FT_BC) {
                if (bi.bct[f] == BC_NEUMANN) {
                    off = T(0);
                    offval = T(0);
                } else {
                    off = wr[f];
                    if (SMOOTH) {
                        offval = T(0);
                        off *= T(2);
                    } else {
                        offval = -off * phic;
                    }
                }
            } else {
                off = wr[f];
                offval = off * tile[ni];
            }
        }
        diag += off;
        offdiag += offval;
    }
    if (SMOOTH) {
        return vfv > 0 ? (rhsv * dv * vfv + offdiag) / diag : T(0);
    } else {
        const T Lphi = (phic * diag - offdiag) / dv;
        return vfv > 0 ? Lphi / vfv : T(0);
    }
}

// Everything one cell needs from HBM that does not depend on the block table: loaded at
// kernel entry, so these loads overlap the table / halo round trips instead of following them.
template <class T>
struct CellIn {
    T x, vf, rhs, dotv;
    T w[6], wcf[6];
};

template <int D, int NB, int MODE, class T>
__device__ inline void load_cell(CellIn<T>& in, const StencilArgsT<T>& a, int b, int c, bool active) {
    using G = Geo<D, NB>;
    int i, j, k;
    G::cell_ijk(c, i, j, k);
    const int gi = b * G::NCELL + c;
    in.x = a.x[gi];
    in.vf = in.rhs = in.dotv = T(0);
#pragma unroll
    for (int f = 0; f < 6; f++) in.w[f] = in.wcf[f] = T(0);
    if (!active) return;
    in.vf = a.vf[gi];
    if (MODE == ST_SMOOTH || MODE == ST_SMOOTH_DB || MODE == ST_RESID || MODE == ST_RESID_INIT) in.rhs = a.rhs[gi];
    if (MODE == ST_APPLY_DOT || MODE == ST_APPLY_DOT2) in.dotv = a.dotv[gi];
#pragma unroll
    for (int f = 0; f < G::NF; f++) {
        const int d = f >> 1;
        in.w[f] = a.w[(b * D + d) * G::NFACE + G::fidx(d, i + (f == 1), j + (f == 3), k + (f == 5))];
    }
}

// The coarse-side weights of the cell's coarse/fine faces (the only faces that use them), once
// the block table is in LDS; issued beside the halo gathers, so they add no round trip.
template <int D, int NB, class T>
__device__ inline void load_wcf(CellIn<T>& in, const StencilArgsT<T>& a, int b, int c, const BlockInfo& bi) {
    using G = Geo<D, NB>;
    int i, j, k;
    G::cell_ijk(c, i, j, k);
#pragma unroll
    for (int f = 0; f < G::NF; f++) {
        const int d = f >> 1, s = f & 1;
        const int idx = d == 0 ? i : (d == 1 ? j : k);
        const int ext = d == 2 ? G::NBZ : NB;
        if ((s ? idx == ext - 1 : idx == 0) && face_type<D, NB>(bi, f, a.m) == FT_CF) {
            int a1, a2;
            tang_axes<D>(d, a1, a2);
            const int ijk[3] = {i, j, k};
            in.wcf[f] = a.wcf[(b * G::NF + f) * G::NSLAB + ijk[a1] + NB * (D == 3 ? ijk[a2] : 0)];
        }
    }
}

constexpr int BI_WORDS = sizeof(BlockInfo) / sizeof(int);

// Operator / residual over one block list (one workgroup per block; the colour passes are
// k_smooth's).  Dependent memory round trips per workgroup: block id -> {table, own cells,
// weights, vf, rhs in parallel} -> halo (needs the table) -> store.  FP64 modes with a fused
// dot product finish with the grid reduction (reference order when rb.exact).
// Threads of a k_stencil / k_smooth workgroup.  On the 139 M-cell 16^3-block snapshot 1 024
// threads ran the FP64 operator at 4.38 against 4.64 TB/s with 512 (80 VGPRs, 6 waves per
// SIMD), and the FP64 colour pass at 2.47 against 2.36 TB/s.
template <int D, int NB>
struct StencilCfg {
    static constexpr int NT = Geo<D, NB>::THREADS;
};
template <int D, int NB>
struct SmoothCfg {
    static constexpr int NT = NB >= 16 ? 1024 : Geo<D, NB>::THREADS;
};

template <int D, int NB, int MODE, class T>
__global__ void __launch_bounds__((StencilCfg<D, NB>::NT))
k_stencil(StencilArgsT<T> a, RedBuf rb, int fin) {
    using G = Geo<D, NB>;
    constexpr int NT = StencilCfg<D, NB>::NT;
    constexpr int CPT = (G::NCELL + NT - 1) / NT;
    constexpr bool REDUCE = MODE == ST_APPLY_DOT || MODE == ST_APPLY_DOT2 || MODE == ST_RESID_INIT;
    static_assert(MODE != ST_SMOOTH && MODE != ST_SMOOTH_DB, "the colour passes are k_smooth's");
    static_assert(!REDUCE || sizeof(T) == 8, "fused reductions are FP64 only");
    // the tile doubles as the reduction's staging (2*min(NCELL, NT)+4 doubles): sized by the
    // larger only when the mode reduces, so that 16^3 blocks keep 3 workgroups per CU
    constexpr int NRED = REDUCE ? 2 * (G::NCELL < NT ? G::NCELL : NT) + 4 : 0;
    __shared__ __attribute__((aligned(16))) T tile[G::NTILE > NRED ? G::NTILE : NRED];
    __shared__ double lds[3 * ((NT + 63) / 64)];
    __shared__ int biw[BI_WORDS];
    const int b = a.list[blockIdx.x];
    if ((int)threadIdx.x < BI_WORDS) biw[threadIdx.x] = reinterpret_cast<const int*>(a.binfo + b)[threadIdx.x];
    // One cell per thread: prefetch everything at entry.  Several cells per thread (16^3
    // blocks): prefetch the own values only, the rest per cell (the registers would spill).
    constexpr bool PRE = CPT == 1;
    CellIn<T> in[1];
    T xo[CPT];
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int c = threadIdx.x + q * NT;
        if (c >= G::NCELL) continue;
        int i, j, k;
        G::cell_ijk(c, i, j, k);
        if (PRE) {
            load_cell<D, NB, MODE, T>(in[0], a, b, c, true);
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
    load_halo<D, NB, T, NT>(tile, a.x, a.w, bi, a.m);
    if (PRE) {
        const int c = threadIdx.x;
        int i, j, k;
        G::cell_ijk(c, i, j, k);
        if (c < G::NCELL) load_wcf<D, NB, T>(in[0], a, b, c, bi);
    }
    __syncthreads();
    double q0[CPT], q1[CPT], mx = 0.0;
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        q0[q] = q1[q] = 0.0;
        const int c = threadIdx.x + q * NT;
        if (c >= G::NCELL) continue;
        int i, j, k;
        G::cell_ijk(c, i, j, k);
        const int gi = b * G::NCELL + c;
        if (!PRE) {
            load_cell<D, NB, MODE, T>(in[0], a, b, c, true);
            load_wcf<D, NB, T>(in[0], a, b, c, bi);
        }
        const T vfv = in[0].vf;
        {
            const T Lp = cell_op<D, NB, false, T>(tile, in[0].w, in[0].wcf, i, j, k, bi, a.m, a.alpha, dv, vfv, T(0));
            if (MODE == ST_APPLY || MODE == ST_APPLY_DOT || MODE == ST_APPLY_DOT2) {
                a.out[gi] = Lp;
                if (MODE == ST_APPLY_DOT) q0[q] = in[0].dotv * Lp;   // rhat . v
                if (MODE == ST_APPLY_DOT2) {
                    q0[q] = Lp * in[0].dotv;                          // t . s
                    q1[q] = Lp * Lp;                                  // t . t
                }
            } else {
                const T r = vfv > 0 ? in[0].rhs - Lp : T(0);
                a.out[gi] = r;
                if (MODE == ST_RESID_INIT) {
                    a.out2[gi] = r;
                    a.out3[gi] = T(0);
                    q0[q] = r * r;                                    // rhat . r with rhat = r
                    mx = fmax(mx, fabs((double)r));
                }
            }
        }
    }
    if constexpr (REDUCE)
        grid_reduce_cells<NT, G::NCELL, CPT>(q0, q1, mx, MODE == ST_APPLY_DOT2 ? 2 : 1, rb,
                                             reinterpret_cast<double*>(tile), lds);
}

// One colour of a red-black Gauss-Seidel pass (MODE ST_SMOOTH in place, ST_SMOOTH_DB from x
// into out with the other colour's cells copied), the per-cell arithmetic of k_stencil's
// smoothing modes (cell_op) with the updates packed by colour.  Thread t updates the colour's
// cells t, t + NT, ... in the order i/2 fastest, then j, then k; with k_stencil's mapping
// (cell t + q NT) every cell of a thread has one parity, so half the lanes of every wave sat
// masked through a colour pass.  The block and its one-cell halo are staged in LDS as there.
template <int NB>
__device__ inline void colour_cell(int a, int colour, int& i, int& j, int& k) {
    constexpr int H = NB / 2;
    j = (a / H) % NB;
    k = a / (H * NB);
    i = 2 * (a % H) + ((colour + j + k) & 1);
}

template <int D, int NB, int MODE, class T>
__global__ void __launch_bounds__((SmoothCfg<D, NB>::NT)) k_smooth(StencilArgsT<T> a) {
    using G = Geo<D, NB>;
    static_assert(MODE == ST_SMOOTH || MODE == ST_SMOOTH_DB, "smoothing modes only");
    static_assert(NB % 2 == 0, "even block edge");
    constexpr int NT = SmoothCfg<D, NB>::NT;
    constexpr int CPT = (G::NCELL + NT - 1) / NT;   // cells per thread (tile fill, copies)
    constexpr int NACT = G::NCELL / 2;               // cells of one colour
    constexpr int APT = (NACT + NT - 1) / NT;        // updates per thread
    constexpr bool PRE = APT == 1;                   // one update per thread: its inputs at entry
    __shared__ __attribute__((aligned(16))) T tile[G::NTILE];
    __shared__ int biw[BI_WORDS];
    const int b = a.list[blockIdx.x];
    if ((int)threadIdx.x < BI_WORDS) biw[threadIdx.x] = reinterpret_cast<const int*>(a.binfo + b)[threadIdx.x];
    T xo[CPT];
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int c = threadIdx.x + q * NT;
        xo[q] = c < G::NCELL ? a.x[b * G::NCELL + c] : T(0);
    }
    CellIn<T> in[1];
    if (PRE) {
        const int ia = threadIdx.x;
        if (ia < NACT) {
            int i, j, k;
            colour_cell<NB>(ia, a.colour, i, j, k);
            load_cell<D, NB, MODE, T>(in[0], a, b, G::cidx(i, j, k), true);
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
        if (MODE == ST_SMOOTH_DB && ((i + j + k) & 1) != a.colour) a.out[b * G::NCELL + c] = xo[q];
    }
    load_halo<D, NB, T, NT>(tile, a.x, a.w, bi, a.m);
    if (PRE) {
        const int ia = threadIdx.x;
        if (ia < NACT) {
            int i, j, k;
            colour_cell<NB>(ia, a.colour, i, j, k);
            load_wcf<D, NB, T>(in[0], a, b, G::cidx(i, j, k), bi);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < APT; q++) {
        const int ia = threadIdx.x + q * NT;
        if (ia >= NACT) continue;
        int i, j, k;
        colour_cell<NB>(ia, a.colour, i, j, k);
        const int c = G::cidx(i, j, k);
        CellIn<T>& ci = in[0];
        if (!PRE) {
            load_cell<D, NB, MODE, T>(ci, a, b, c, true);
            load_wcf<D, NB, T>(ci, a, b, c, bi);
        }
        const T u = cell_op<D, NB, true, T>(tile, ci.w, ci.wcf, i, j, k, bi, a.m, a.alpha, dv, ci.vf, ci.rhs);
        if (MODE == ST_SMOOTH_DB) a.out[b * G::NCELL + c] = u;
        else const_cast<T*>(a.x)[b * G::NCELL + c] = u;
    }
}

// --------------------------------------------------------------- restriction / prolongation
// mg_restict_wgt onto the non-leaf blocks of one level (SayakaMacIB.cpp:157-232 with
// restrictCellBlockData_Average, SayakaInterpPatch.cpp:398-444): coarse =
// ((sum over the 2^d children, kk,jj,ii order, of x*vf) / 2^d) / vf_coarse, 0 if vf_coarse=0.
template <int D, int NB>
__global__ void __launch_bounds__((Geo<D, NB>::THREADS))
k_restrict(const int* list, const BlockInfo* binfo, const double* vf, double* x) {
    using G = Geo<D, NB>;
    const int b = list[blockIdx.x];
    const double wgt = 1.0 / G::NCHILD;
    for (int c = threadIdx.x; c < G::NCELL; c += G::THREADS) {
        int i, j, k;
        G::cell_ijk(c, i, j, k);
        const int h = NB / 2;
        const int si = i >= h, sj = j >= h, sk = (D == 3 && k >= h);
        const int ch = si | (sj << 1) | (sk << 2);
        const int fb = binfo[b].child[ch];
        const int fi0 = (i - si * h) * 2, fj0 = (j - sj * h) * 2, fk0 = D == 3 ? (k - sk * h) * 2 : 0;
        double sum = 0.0;
        for (int kk = 0; kk <= (D == 3 ? 1 : 0); kk++)
            for (int jj = 0; jj <= 1; jj++)
                for (int ii = 0; ii <= 1; ii++) {
                    const int fi = fb * G::NCELL + G::cidx(fi0 + ii, fj0 + jj, fk0 + kk);
                    sum += x[fi] * vf[fi];
                }
        const double v = vf[b * G::NCELL + c];
        const double r = sum * wgt;
        x[b * G::NCELL + c] = v > 0 ? r / v : 0.0;
    }
}

// mg_prolong_wgt on every block of one level (SayakaMacIB.cpp:234-283): injection from the
// parent cell, zero where vf = 0.
template <int D, int NB>
__global__ void __launch_bounds__((Geo<D, NB>::THREADS))
k_prolong(const int* list, const BlockInfo* binfo, const double* vf, double* x) {
    using G = Geo<D, NB>;
    const int b = list[blockIdx.x];
    const int p = binfo[b].parent;
    const int wc = binfo[b].which_child;
    const int h = NB / 2;
    for (int c = threadIdx.x; c < G::NCELL; c += G::THREADS) {
        int i, j, k;
        G::cell_ijk(c, i, j, k);
        const int pc = G::cidx((i >> 1) + side_of(wc, 0) * h, (j >> 1) + side_of(wc, 1) * h,
                               D == 3 ? (k >> 1) + side_of(wc, 2) * h : 0);
        const int gi = b * G::NCELL + c;
        x[gi] = vf[gi] > 0 ? x[p * G::NCELL + pc] : 0.0;
    }
}

// --------------------------------------------------------------------------- vector kernels
enum VecOp {
    V_PINIT = 0,   // p = r ; resid = p masked ; (phat = p if no preconditioner)
    V_PUPDATE,     // p = r + beta*(p - omega*v) ; resid = p masked ; (phat = p)
    V_SUPDATE,     // s = r - alpha*v ; x = x + alpha*phat ; resid = s masked ; |s| ; (shat = s)
    V_RUPDATE,     // x = x + omega*shat ; r = s - omega*t ; |r| ; rhat.r
    V_FINAL,       // sol = sol + x
    V_COPYPLUS0,   // dst = 0.0 + src   (SOL = 0 + CORR, then phat = SOL)
    V_MASKCOPY,    // dst = vf>0 ? src : 0
    V_ZERO,        // dst = 0
    V_DOT,         // dot(a, b), max|a|
};

struct VecArgs {
    const int* list;
    const double* vf;
    const double* r;
    double* p;
    const double* v;
    double* resid;
    double* phat;     // when non-null and no preconditioner, receives a copy
    double* s;
    double* x;
    const double* t;
    const double* rhat;
    const double* src;
    double* dst;
    const double* scal;
    const int* xskip;   // V_SUPDATE: x is left as it is when *xskip != 0 (a queued half step of a
                        // solve that has stopped)
};

template <int NCELL, int NT>
__global__ void __launch_bounds__(NT) k_vec(VecArgs a, int op, RedBuf rb, int fin) {
    constexpr int CPT = (NCELL + NT - 1) / NT;
    __shared__ double lds[3 * ((NT + 63) / 64)];
    __shared__ __attribute__((aligned(16))) double scratch[2 * (NCELL < NT ? NCELL : NT) + 4];
    const int b = a.list[blockIdx.x];
    double q0[CPT], q1[CPT], mx = 0.0;
    const double* S = a.scal;
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        q0[q] = q1[q] = 0.0;
        const int c = threadIdx.x + q * NT;
        if (c >= NCELL) continue;
        const int gi = b * NCELL + c;
        switch (op) {
            case V_PINIT: {
                const double pv = a.r[gi];
                a.p[gi] = pv;
                a.resid[gi] = a.vf[gi] > 0 ? pv : 0.0;
                if (a.phat) a.phat[gi] = pv;
                break;
            }
            case V_PUPDATE: {
                // beta = (rho/rho1)*(alpha/omega); p = p - omega*v; p = r + beta*p
                // (SayakaMacSolver.cpp:782-790, Calc_axpby SayakaBlockData.cpp:253-255)
                const double beta = (S[S_RHO] / S[S_RHO1]) * (S[S_ALPHA] / S[S_OMEGA]);
                const double p1 = 1.0 * a.p[gi] + -S[S_OMEGA] * a.v[gi];
                const double pv = 1.0 * a.r[gi] + beta * p1;
                a.p[gi] = pv;
                a.resid[gi] = a.vf[gi] > 0 ? pv : 0.0;
                if (a.phat) a.phat[gi] = pv;
                break;
            }
            case V_SUPDATE: {
                const double al = S[S_ALPHA];
                const double sv = 1.0 * a.r[gi] + -al * a.v[gi];
                a.s[gi] = sv;
                if (!a.xskip || !*a.xskip) a.x[gi] = 1.0 * a.x[gi] + al * a.phat[gi];
                if (a.resid) a.resid[gi] = a.vf[gi] > 0 ? sv : 0.0;
                if (a.dst) a.dst[gi] = sv;
                mx = fmax(mx, fabs(sv));
                break;
            }
            case V_RUPDATE: {   // x (into dst when given) += omega shat ; r = s - omega t
                const double om = S[S_OMEGA];
                (a.dst ? a.dst : a.x)[gi] = 1.0 * a.x[gi] + om * a.src[gi];
                const double rv = 1.0 * a.s[gi] + -om * a.t[gi];
                const_cast<double*>(a.r)[gi] = rv;
                mx = fmax(mx, fabs(rv));
                q0[q] = a.rhat[gi] * rv;
                break;
            }
            case V_FINAL: a.dst[gi] = a.dst[gi] + a.x[gi]; break;
            case V_COPYPLUS0: a.dst[gi] = 0.0 + a.src[gi]; break;
            case V_MASKCOPY: a.dst[gi] = a.vf[gi] > 0 ? a.src[gi] : 0.0; break;
            case V_ZERO: a.dst[gi] = 0.0; break;
            case V_DOT: {
                const double u = a.src[gi];
                q0[q] = u * a.t[gi];
                mx = fmax(mx, fabs(u));
                break;
            }
            default: break;
        }
    }
    if (fin != FIN_NONE)
        grid_reduce_cells<NT, NCELL, CPT>(q0, q1, mx, (op == V_RUPDATE || op == V_DOT) ? 1 : 0, rb, scratch, lds);
}

}  // namespace sydev
