// sayaka_dev.hip -- device context, index tables and the C ABI of include/sayaka_amd.h.
//
// The host keeps the tree; after each regrid it hands the per-node arrays to
// sayaka_dev_set_topology, which builds the per-block index table (BlockInfo), the block
// metric and the block lists every launch iterates over:
//   level l, all blocks / leaves / non-leaves (ascending id, SayakaRegrid.cpp:962-976);
//   tower T(m) = leaves of levels < m and every block of level m, ordered m..1
//   (MGLevelTower::update, SayakaTower.cpp:28-68), with one sub-list per level.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <chrono>
#include <thread>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/sayaka_amd.h"
#include "bottom.h"
#include "coef.h"
#include "cnn.h"
#include "cnn_data.h"
#include "kernels.h"
#include "mixed.h"
#include "flux.h"
#include "remap.h"
#include "projection.h"
#include "statefill.h"

using namespace sydev;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return fail(-3, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int MAXLV = 16;

// Allocation-free tower dot product (sayaka_dev_dot_noalloc): its per-workgroup partials and
// scalars live in this file-scope array, one copy per GPU, held until the process ends.  Host
// code must not pass the array's name to a kernel (that is a host-side stand-in, not a device
// address): the device address of each GPU's copy is looked up once, at context creation
// with that GPU current, and kept in scratch_dev[device].
constexpr int NOALLOC_MAX_BLOCKS = 1 << 16;                       // blocks of one tower
constexpr size_t SCRATCH_DOUBLES = 3 * (size_t)NOALLOC_MAX_BLOCKS + 64;   // partials, then scalars
__device__ __align__(256) double scratch[SCRATCH_DOUBLES];
constexpr int SCRATCH_MAX_DEV = 64;
std::mutex scratch_mu;
double* scratch_dev[SCRATCH_MAX_DEV] = {};

// device address of this GPU's scratch array (the calling thread has made `device` current)
hipError_t scratch_lookup(int device) {
    if (device < 0 || device >= SCRATCH_MAX_DEV) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(scratch_mu);
    if (scratch_dev[device]) return hipSuccess;
    double* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e == hipSuccess) scratch_dev[device] = ptr;
    return e;
}

struct ListRef {
    int off = 0, n = 0;
};

}  // namespace

struct sayaka_dev_ctx {
    int device = 0, ndim = 3, nb = 8, cap = 0;
    int ncell = 0, nface = 0, nslab = 0, nf = 0;
    int nblk = 0, minlevel = 1, maxlevel = 1, finest = 1;
    hipStream_t stream = nullptr;
    double alpha = 0.0, beta = 1.0;

    std::vector<BlockInfo> h_binfo;
    std::vector<BlockMetric> h_met;
    BlockInfo* d_binfo = nullptr;
    BlockMetric* d_met = nullptr;

    std::vector<int> h_lists;
    int* d_lists = nullptr;
    ListRef lv_all[MAXLV + 1], lv_leaf[MAXLV + 1], lv_nonleaf[MAXLV + 1];
    ListRef tw[MAXLV + 1];                    // whole tower, sub-levels m..1
    ListRef tw_sub[MAXLV + 1][MAXLV + 1];     // [m][l]
    ListRef all_blocks;
    ListRef by_level;                          // every block, levels 1.. in turn, ascending id

    double* d_field[SAYAKA_NFIELD] = {};
    double* d_vf = nullptr;
    double* d_af = nullptr;   // [blk][D][NFACE]
    double* d_b = nullptr;
    double* d_w = nullptr;
    double* d_wcf = nullptr;

    double* d_partials = nullptr;
    double* d_scal = nullptr;
    double* h_scal = nullptr;  // pinned
    double* d_bot_vpub = nullptr;
    double* d_bot_rpub = nullptr;
    double* d_bot_part = nullptr;
    int* d_bot_stats = nullptr;
    // Stopping tests on the device (RedBuf::stop), for the half-iterations a solve queues before
    // it has read whether the solve stops: while stop_flags is set, the FIN_SNORM / FIN_RUPDATE
    // finishes write d_stop[0] / d_stop[1]; the coarsest-level solves launched while skip_ptr is
    // set do nothing when *skip_ptr is 1.
    int* d_stop = nullptr;
    bool stop_flags = false;
    const int* skip_ptr = nullptr;
    double stop_tol_rel = 0.0, stop_tol_abs = 0.0;
    double defer_factor = 100.0;   // sayaka_dev_set_half_step_defer
    // FIN_RUPDATE finishes launched while publish is set copy the scalars into coherent pinned
    // memory: h_mirror[0 .. S_NSCALAR) and the sequence number in h_mirror[S_NSCALAR] (d_mirror =
    // its device address); the host reads them without waiting for the rest of the stream.
    double* h_mirror = nullptr;
    double* d_mirror = nullptr;
    bool publish = false;
    unsigned long long pub_seq = 0;
    unsigned long long* d_bot_flags = nullptr;   // grid-barrier flags of the bottom solve, one per workgroup
    unsigned long long bot_launches = 0;
    int bot_fit[2] = {0, 0};                     // workgroups of k_bottom<.., PC> resident at once
    int bot_fit_xcd[2] = {0, 0};                 // ... on one XCD
    int bot_one_xcd = 1;                         // place the coarsest solve on one XCD when it fits
    long long bot_place[3] = {0, 0, 0};          // launches; on one XCD; of those, fallen back
    int* h_bot_stats = nullptr;  // pinned

    MixedState mixed;
    // learned initial guess (8^3 blocks in 3D): parameters, activations, training store
    struct Cnn {
        int cap = 0;                      // windows per forward / backward pass
        float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
        __bf16* wf[cnn::NLAYER] = {};
        __bf16* wb[cnn::NLAYER] = {};
        __bf16* act[cnn::NLAYER] = {};    // act[0] input windows, act[l] output of layer l
        __bf16 *g0 = nullptr, *g1 = nullptr;   // activation gradients (ping-pong)
        float* out32 = nullptr;
        float* tgt = nullptr;
        float* loss = nullptr;
        int* list = nullptr;              // leaf blocks of the current batch / guess
        int list_cap = 0;
        __bf16* store = nullptr;          // training windows [n][NV][C]
        float* store_t = nullptr;         // training targets [n][512]
        long long store_n = 0, store_cap = 0;
        int step = 0;
    } cnn;
    double* d_gen = nullptr;              // generator scratch: block corners, sizes, partial sums
    int gen_cap = 0;
    // per-kernel-class launch timing (HIP events on the context stream)
    struct ProfRec {
        hipEvent_t a, b;
        long long cells;
        double bytes;
    };
    bool prof_on[SAYAKA_KCLASS_N] = {};
    std::vector<ProfRec> prof[SAYAKA_KCLASS_N];
    std::vector<hipEvent_t> ev_pool;
    int coef_version = 0;
    // set_topology leaves the coefficients unit (the reference's prepareSolver) without
    // computing them: the first entry point that reads them fills and derives them
    // (ensure_coefs), unless set_coefs or the generators write their own first
    int coef_pending = 0;
    int exact = 1;  // reference-order reductions (FP64 solves); the mixed solve clears it
    // two-level preconditioner of the mixed solve's coarsest problem (bottom.h)
    int* d_l1pos = nullptr;        // block id -> level-1 list position, -1 elsewhere
    double* d_pc_jac = nullptr;    // [PC_MAXBLK][ncell]
    double* d_pc_fsum = nullptr;   // [PC_MAXBLK][6]
    float* d_pc_s0inv = nullptr;   // [PC_MAXBLK][PC_MAXBLK] (row-major nl x nl in use)
    int pc_enable = 1, pc_version = -1, pc_ok = 0;
    double* d_flux = nullptr;      // face flux of one direction, [cap][nface] (allocated on first use)
    double* d_ghost = nullptr;     // one-layer face ghosts, [cap][2 ndim][nslab] (allocated on first use)
    // regrid data movement (allocated on first use): old field copy, origin map, new-block list
    double* d_remap_tmp = nullptr;
    int* d_remap_idx = nullptr;
    // projection (allocated on first use): face velocity [cap][D][nface], divergence [cap][ncell]
    double* d_vel = nullptr;
    double* d_div = nullptr;
    int vel_valid = 0;             // d_vel generated for the current tree
    // general state fill (statefill.h): the tree's surrounding tables [blk][27] (host copy kept
    // to check BC maps) and the BC tables of the last call
    int* d_surr = nullptr;
    int* d_bcref = nullptr;
    std::vector<int> h_surr;
    int surr_n = 0;
    int* d_fill_int = nullptr;     // physical codes, then [nc][27][nphys] math BC types
    double* d_fill_val = nullptr;  // [nc][27][nphys] BC values
    size_t fill_icap = 0, fill_vcap = 0;
    int* d_remap = nullptr;        // state remap: origin, then the per-level block lists
    size_t remap_cap = 0;
    int* d_move = nullptr;         // block move: origin per slot
    size_t move_cap = 0;
    // host <-> device traffic: state bytes each way, topology and surroundings uploads
    long long xfer[4] = {0, 0, 0, 0};
};

namespace {

void cnn_free(sayaka_dev_ctx::Cnn& n);

constexpr int NUM_XCD = 8;   // MI355X: 8 XCDs of 32 CUs, each with its own L2

// The coarsest solve's launch counters of the last readback (h_bot_stats) into the context's
// running totals.
void note_bottom_stats(sayaka_dev_ctx* c) {
    c->bot_place[0] += c->h_bot_stats[0];
    c->bot_place[1] += c->h_bot_stats[5];
    c->bot_place[2] += c->h_bot_stats[6];
}

// Streaming device copy (TreeData::Copy of a whole field): one 16-byte element per lane and a
// grid that covers the field.  Measured on MI355X over 2 GiB: 6.2 TB/s (reads + writes), against
// 4.5-5.2 TB/s for grid-stride loops of 1 024-8 192 workgroups and 5.0 TB/s for hipMemcpy.
typedef double v2d __attribute__((ext_vector_type(2)));
__global__ void __launch_bounds__(256) k_copy16(const v2d* __restrict__ src, v2d* __restrict__ dst, size_t n2) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n2) dst[i] = src[i];
}

__global__ void k_copy_dot_norm(const double* __restrict__ scal, double* __restrict__ out2) {
    if (threadIdx.x == 0) {
        out2[0] = scal[S_DOT];
        out2[1] = scal[S_NORM];
    }
}

// The regrid's packing of a caller's state: slot b of every component takes slot origin[b]
// (zeros where origin[b] == -1 and for slots b >= n).  One workgroup per (slot, component).
__global__ void __launch_bounds__(256) k_block_move(const double* __restrict__ src, double* __restrict__ dst,
                                                    const int* __restrict__ origin, long long comp_stride,
                                                    long long block_len, int n) {
    const int b = blockIdx.x, comp = blockIdx.y;
    const int o = b < n ? origin[b] : -1;
    double* d = dst + comp * comp_stride + (long long)b * block_len;
    if (o < 0) {
        for (long long q = threadIdx.x; q < block_len; q += blockDim.x) d[q] = 0.0;
        return;
    }
    const double* s = src + comp * comp_stride + (long long)o * block_len;
    for (long long q = threadIdx.x; q < block_len; q += blockDim.x) d[q] = s[q];
}

__global__ void k_fill(double* __restrict__ p, size_t n, double v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

hipEvent_t pool_event(sayaka_dev_ctx* c) {
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Brackets one launch of kernel class k with events when timing of that class is on;
// cells = cells the launch covers, bytes = its algorithmic HBM bytes (below).
struct ProfScope {
    sayaka_dev_ctx* c;
    int k;
    long long cells;
    double bytes;
    bool kernel_timed;
    hipEvent_t a = nullptr, b = nullptr;
    // kernel_timed: the launch records the two events itself (hipExtLaunchKernel's start / stop
    // events, the dispatch's own timestamps, which profilers read too); otherwise the scope
    // records them around whatever it brackets (a marker pair adds a few microseconds)
    ProfScope(sayaka_dev_ctx* c_, int k_, long long cells_, double bytes_, bool kernel_timed_ = false)
        : c(c_), k(k_), cells(cells_), bytes(bytes_), kernel_timed(kernel_timed_) {
        if (c->prof_on[k]) {
            a = pool_event(c);
            if (a && kernel_timed) b = pool_event(c);
            if (a && !kernel_timed) hipEventRecord(a, c->stream);
        }
    }
    hipEvent_t start() const { return kernel_timed ? a : nullptr; }
    hipEvent_t stop() const { return kernel_timed ? b : nullptr; }
    ~ProfScope() {
        if (!a) return;
        if (!kernel_timed) {
            b = pool_event(c);
            if (!b) return;
            hipEventRecord(b, c->stream);
        }
        if (!b) return;
        c->prof[k].push_back({a, b, cells, bytes});
    }
};

// Algorithmic HBM bytes per block of one launch: every value the operation must read or
// write once (valid cells; face weights at (nb+1)/nb faces per cell and direction; ghosts,
// index tables and reduction partials excluded).  C = one cell field, Fw = the weights.
template <int D, int NB, class T>
struct AlgBytes {
    static constexpr double C = (double)Geo<D, NB>::NCELL * sizeof(T);
    static constexpr double Fw = (double)D * Geo<D, NB>::NFACE * sizeof(T);
    static double stencil(int mode) {
        switch (mode) {
            case ST_APPLY: return 3 * C + Fw;                // x, vf, w -> out
            case ST_RESID: return 4 * C + Fw;                // + rhs
            case ST_SMOOTH: return 0.5 * (4 * C + Fw);       // half of a full sweep (x rw, rhs, vf, w)
            case ST_APPLY_DOT: return 4 * C + Fw;            // apply + rhat
            case ST_APPLY_DOT2: return 4 * C + Fw;           // apply + s
            case ST_RESID_INIT: return 6 * C + Fw;           // resid -> r, rhat, x = 0
            default: return 0.0;
        }
    }
    static double vec(int op, bool extra) {
        switch (op) {
            case V_PINIT: return (extra ? 5 : 4) * C;        // r, vf -> p, resid (, phat)
            case V_PUPDATE: return (extra ? 7 : 6) * C;      // r, p, v, vf -> p, resid (, phat)
            case V_SUPDATE: return (extra ? 9 : 8) * C;      // r, v, x, phat, vf -> s, x, resid (, shat)
            case V_RUPDATE: return 7 * C;                    // x, shat, s, t, rhat -> x, r
            case V_FINAL: return 3 * C;
            case V_COPYPLUS0: return 2 * C;
            case V_MASKCOPY: return 3 * C;
            case V_ZERO: return C;
            case V_DOT: return 2 * C;
            default: return 0.0;
        }
    }
    // per coarse (parent) block: 2^d children read x and vf, parent vf read and x written
    static constexpr double restrict_() { return (1 << D) * 2 * C + 2 * C; }
    // per fine block: parent value (1/2^d of a field), vf, write
    static constexpr double prolong() { return 2 * C + C / (1 << D); }
};

RedBuf redbuf(sayaka_dev_ctx* c) { return RedBuf{c->d_partials, c->d_scal, c->exact}; }

const int* dlist(sayaka_dev_ctx* c, ListRef r) { return c->d_lists + r.off; }

StencilArgs stencil_args(sayaka_dev_ctx* c, ListRef lst, int m) {
    StencilArgs a{};
    a.list = dlist(c, lst);
    a.nlist = lst.n;
    a.m = m;
    a.binfo = c->d_binfo;
    a.met = c->d_met;
    a.w = c->d_w;
    a.wcf = c->d_wcf;
    a.vf = c->d_vf;
    a.alpha = c->alpha;
    return a;
}

template <int D, int NB>
int launch_stencil(sayaka_dev_ctx* c, int mode, const StencilArgs& a, int fin) {
    using G = Geo<D, NB>;
    if (a.nlist <= 0) return 0;
    dim3 grid(a.nlist), block(StencilCfg<D, NB>::NT);
    RedBuf rb = redbuf(c);
    ProfScope ps(c, mode == ST_SMOOTH ? SAYAKA_KCLASS_SMOOTH : SAYAKA_KCLASS_APPLY, (long long)a.nlist * G::NCELL,
                 a.nlist * AlgBytes<D, NB, double>::stencil(mode), mode == ST_SMOOTH);
    switch (mode) {
        case ST_APPLY: hipLaunchKernelGGL((k_stencil<D, NB, ST_APPLY, double>), grid, block, 0, c->stream, a, rb, fin); break;
        case ST_RESID: hipLaunchKernelGGL((k_stencil<D, NB, ST_RESID, double>), grid, block, 0, c->stream, a, rb, fin); break;
        case ST_SMOOTH:
            hipExtLaunchKernelGGL((k_smooth<D, NB, ST_SMOOTH, double>), grid, dim3(SmoothCfg<D, NB>::NT), 0,
                                  c->stream, ps.start(), ps.stop(), 0, a);
            break;
        case ST_APPLY_DOT: hipLaunchKernelGGL((k_stencil<D, NB, ST_APPLY_DOT, double>), grid, block, 0, c->stream, a, rb, fin); break;
        case ST_APPLY_DOT2: hipLaunchKernelGGL((k_stencil<D, NB, ST_APPLY_DOT2, double>), grid, block, 0, c->stream, a, rb, fin); break;
        case ST_RESID_INIT: hipLaunchKernelGGL((k_stencil<D, NB, ST_RESID_INIT, double>), grid, block, 0, c->stream, a, rb, fin); break;
        default: return fail(-2, "bad stencil mode");
    }
    if (mode == ST_APPLY_DOT || mode == ST_APPLY_DOT2 || mode == ST_RESID_INIT)
        hipLaunchKernelGGL(k_red_finish, dim3(1), dim3(RED_NT), 0, c->stream, rb, a.nlist,
                           mode == ST_APPLY_DOT2 ? 2 : 1, fin);
    HIPCHK(hipGetLastError());
    return 0;
}

template <int D, int NB>
int launch_vec(sayaka_dev_ctx* c, ListRef lst, int op, const VecArgs& a, int fin) {
    using G = Geo<D, NB>;
    if (lst.n <= 0) return 0;
    const bool extra = (op == V_PINIT || op == V_PUPDATE) ? a.phat != nullptr : (op == V_SUPDATE && a.dst != nullptr);
    ProfScope ps(c, SAYAKA_KCLASS_VEC, (long long)lst.n * G::NCELL, lst.n * AlgBytes<D, NB, double>::vec(op, extra));
    hipLaunchKernelGGL((k_vec<G::NCELL, G::THREADS>), dim3(lst.n), dim3(G::THREADS), 0, c->stream, a, op,
                       redbuf(c), fin);
    if (fin != FIN_NONE) {
        RedBuf rb = redbuf(c);
        if ((fin == FIN_SNORM || fin == FIN_RUPDATE) && c->stop_flags) {
            rb.stop = c->d_stop;
            rb.tol_rel = c->stop_tol_rel;
            rb.tol_abs = c->stop_tol_abs;
        }
        if (fin == FIN_RUPDATE && c->publish && c->d_mirror) {
            rb.host = c->d_mirror;
            rb.seq = ++c->pub_seq;
        }
        hipLaunchKernelGGL(k_red_finish, dim3(1), dim3(RED_NT), 0, c->stream, rb, lst.n,
                           (op == V_RUPDATE || op == V_DOT) ? 1 : 0, fin);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

VecArgs vec_args(sayaka_dev_ctx* c, ListRef lst) {
    VecArgs a{};
    a.list = dlist(c, lst);
    a.vf = c->d_vf;
    a.scal = c->d_scal;
    return a;
}

// ------------------------------------------------------------------- dimension dispatch
template <int D, int NB>
struct Impl {
    using G = Geo<D, NB>;
    static constexpr int D_ = D, NB_ = NB;

    static int apply(sayaka_dev_ctx* c, int m, int out, int in) {
        StencilArgs a = stencil_args(c, c->tw[m], m);
        a.x = c->d_field[in];
        a.out = c->d_field[out];
        return launch_stencil<D, NB>(c, ST_APPLY, a, FIN_NONE);
    }
    static int residual(sayaka_dev_ctx* c, int m, int r, int x, int rhs) {
        StencilArgs a = stencil_args(c, c->tw[m], m);
        a.x = c->d_field[x];
        a.out = c->d_field[r];
        a.rhs = c->d_field[rhs];
        return launch_stencil<D, NB>(c, ST_RESID, a, FIN_NONE);
    }
    // One mg_smooth_level pass on tower m (SayakaMacOperator.cpp:1028-1112): sub-levels m..1,
    // stop at the first empty one, red then black, each colour one launch.
    static int smooth(sayaka_dev_ctx* c, int m, double* x, const double* rhs) {
        for (int l = m; l >= 1; l--) {
            ListRef sub = c->tw_sub[m][l];
            if (sub.n <= 0) break;
            for (int rb = 0; rb < 2; rb++) {
                StencilArgs a = stencil_args(c, sub, m);
                a.x = x;
                a.rhs = rhs;
                a.colour = rb;
                int e = launch_stencil<D, NB>(c, ST_SMOOTH, a, FIN_NONE);
                if (e) return e;
            }
        }
        return 0;
    }
    static int restrict_(sayaka_dev_ctx* c, int cl, double* x) {
        if (cl >= c->maxlevel || cl < 1) return 0;
        ListRef lst = c->lv_nonleaf[cl];
        if (lst.n <= 0 || c->lv_all[cl + 1].n <= 0) return 0;
        ProfScope ps(c, SAYAKA_KCLASS_TRANSFER, (long long)lst.n * G::NCELL * G::NCHILD,
                     lst.n * AlgBytes<D, NB, double>::restrict_());
        hipLaunchKernelGGL((k_restrict<D, NB>), dim3(lst.n), dim3(G::THREADS), 0, c->stream, dlist(c, lst),
                           c->d_binfo, c->d_vf, x);
        HIPCHK(hipGetLastError());
        return 0;
    }
    static int prolong(sayaka_dev_ctx* c, int fl, double* x) {
        if (fl <= 1 || fl > c->maxlevel) return 0;
        ListRef lst = c->lv_all[fl];
        if (lst.n <= 0) return 0;
        ProfScope ps(c, SAYAKA_KCLASS_TRANSFER, (long long)lst.n * G::NCELL, lst.n * AlgBytes<D, NB, double>::prolong());
        hipLaunchKernelGGL((k_prolong<D, NB>), dim3(lst.n), dim3(G::THREADS), 0, c->stream, dlist(c, lst),
                           c->d_binfo, c->d_vf, x);
        HIPCHK(hipGetLastError());
        return 0;
    }
    static int bottom(sayaka_dev_ctx* c, double* sol, const double* rhs, double tol_rel, double tol_abs, int max_iter,
                      bool two_level = false) {
        ListRef lst = c->tw[1];
        if (lst.n <= 0) return 0;
        two_level = two_level && lst.n <= PC_MAXBLK;
        BottomArgs A{};
        A.st = stencil_args(c, lst, 1);
        A.rhs = rhs;
        A.sol = sol;
        A.vpub = c->d_bot_vpub;
        A.rpub = c->d_bot_rpub;
        A.partials = c->d_bot_part;
        A.stats = c->d_bot_stats;
        A.flags = c->d_bot_flags;
        A.episode0 = (++c->bot_launches) << 20;   // at most 1 + 3 * 2000 barriers per launch
        A.tol_rel = tol_rel;
        A.tol_abs = tol_abs;
        A.max_iter = max_iter;
        A.exact = two_level ? 0 : c->exact;
        A.s0inv = c->d_pc_s0inv;
        A.jac = c->d_pc_jac;
        A.l1pos = c->d_l1pos;
        A.skip = c->skip_ptr;
        using BC = BotCfg<D, NB>;
        const int nwg = (lst.n + BC::BPW - 1) / BC::BPW;
        const void* kern = two_level ? (const void*)k_bottom<D, NB, true> : (const void*)k_bottom<D, NB, false>;
        // The kernel's grid barriers need every workgroup resident at once.  An ordinary launch
        // is enough for that when the grid fits the chip at the kernel's occupancy (checked
        // here; kernels earlier on the stream have finished before it starts); the barrier's
        // bounded wait turns a workgroup that is not scheduled (another process filling the
        // GPU) into error -5 instead of a hang.  A cooperative launch would add nothing but a
        // device-wide lock that keeps a second process from solving on the same GPU.
        int& fit = c->bot_fit[two_level ? 1 : 0];
        int& fit_xcd = c->bot_fit_xcd[two_level ? 1 : 0];
        if (fit == 0) {
            int per_cu = 0, ncu = 0;
            HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, BC::NT, 0));
            HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
            fit = per_cu * ncu;
            fit_xcd = per_cu * (ncu / NUM_XCD);
        }
        if (nwg > fit)
            return fail(-2, "coarsest level too large for one resident grid: " + std::to_string(nwg) +
                                " workgroups, " + std::to_string(fit) + " fit");
        // One XCD when the workers fit there: the dispatcher deals workgroups to the XCDs in
        // turn (workgroup i to XCD i mod 8), so a grid of 8 x nwg puts the workers
        // blockIdx.x % 8 == 0 on one XCD and the others exit at once.  The kernel checks the
        // placement at its first barrier and keeps the device-wide barrier if it does not hold.
        A.confine = c->bot_one_xcd && nwg <= fit_xcd ? 1 : 0;
        const dim3 grid(A.confine ? NUM_XCD * nwg : nwg);
        // iterative: its algorithmic bytes depend on the iteration count, which the caller
        // reads back (bytes recorded as 0); timed by the launch's own start / stop events
        ProfScope ps(c, SAYAKA_KCLASS_BOTTOM, (long long)lst.n * G::NCELL, 0.0, true);
        if (two_level)
            hipExtLaunchKernelGGL((k_bottom<D, NB, true>), grid, dim3(BC::NT), 0, c->stream, ps.start(),
                                  ps.stop(), 0, A);
        else
            hipExtLaunchKernelGGL((k_bottom<D, NB, false>), grid, dim3(BC::NT), 0, c->stream, ps.start(),
                                  ps.stop(), 0, A);
        HIPCHK(hipGetLastError());
        return 0;
    }
    // Two-level preconditioner data for the current coefficients and BCs (bottom.h): the
    // per-cell 1 / diag(A) and per-face weight sums on the device, then the coarse matrix
    // S0 = P^T (D A) P of the level-1 blocks assembled and inverted here (at most 64 x 64).
    // The pure-Neumann S0 is singular (constants); its inverse is taken of S0 + (tr S0 / n^2)
    // 1 1^T, which acts as a pseudo-inverse up to an added constant that A annihilates.
    static int pc_prepare(sayaka_dev_ctx* c) {
        ListRef lst = c->tw[1];
        c->pc_ok = 0;
        if (!c->pc_enable || lst.n <= 0 || lst.n > PC_MAXBLK) return 0;
        const int nl = lst.n;
        StencilArgs a = stencil_args(c, lst, 1);
        hipLaunchKernelGGL((k_l1_coarse<D, NB>), dim3(nl), dim3(G::THREADS), 0, c->stream, a, c->d_pc_jac,
                           c->d_pc_fsum);
        HIPCHK(hipGetLastError());
        std::vector<double> fs((size_t)nl * 6);
        HIPCHK(hipMemcpyAsync(fs.data(), c->d_pc_fsum, fs.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        const int* L = c->h_lists.data() + lst.off;
        std::vector<int> pos(c->nblk, -1);
        for (int q = 0; q < nl; q++) pos[L[q]] = q;
        std::vector<double> S((size_t)nl * nl, 0.0);
        bool singular = true;
        for (int q = 0; q < nl; q++) {
            const BlockInfo& bi = c->h_binfo[L[q]];
            for (int f = 0; f < c->nf; f++) {
                const double w = fs[(size_t)q * 6 + f];
                const int n = bi.nbr[f];
                if (n >= 0 && pos[n] >= 0) {
                    S[(size_t)q * nl + q] += w;
                    S[(size_t)q * nl + pos[n]] -= w;
                } else if (n <= PHYS_BNDRY) {
                    S[(size_t)q * nl + q] += w;   // Dirichlet / SimpleFill wall (0 on Neumann walls)
                    if (w > 0) singular = false;
                }
            }
        }
        double tr = 0.0;
        for (int q = 0; q < nl; q++)
            for (int r = q + 1; r < nl; r++) {
                const double m = 0.5 * (S[(size_t)q * nl + r] + S[(size_t)r * nl + q]);
                S[(size_t)q * nl + r] = S[(size_t)r * nl + q] = m;
            }
        for (int q = 0; q < nl; q++) {
            if (S[(size_t)q * nl + q] == 0.0) S[(size_t)q * nl + q] = 1.0;   // fully solid block
            tr += S[(size_t)q * nl + q];
        }
        if (singular)
            for (auto& v : S) v += tr / ((double)nl * nl);
        // Gauss-Jordan with partial pivoting
        std::vector<double> I((size_t)nl * nl, 0.0);
        for (int q = 0; q < nl; q++) I[(size_t)q * nl + q] = 1.0;
        for (int col = 0; col < nl; col++) {
            int piv = col;
            for (int r = col + 1; r < nl; r++)
                if (std::fabs(S[(size_t)r * nl + col]) > std::fabs(S[(size_t)piv * nl + col])) piv = r;
            if (S[(size_t)piv * nl + col] == 0.0) return 0;   // no preconditioner (plain coarse solve)
            if (piv != col)
                for (int k = 0; k < nl; k++) {
                    std::swap(S[(size_t)piv * nl + k], S[(size_t)col * nl + k]);
                    std::swap(I[(size_t)piv * nl + k], I[(size_t)col * nl + k]);
                }
            const double d = 1.0 / S[(size_t)col * nl + col];
            for (int k = 0; k < nl; k++) {
                S[(size_t)col * nl + k] *= d;
                I[(size_t)col * nl + k] *= d;
            }
            for (int r = 0; r < nl; r++) {
                if (r == col) continue;
                const double m = S[(size_t)r * nl + col];
                if (m == 0.0) continue;
                for (int k = 0; k < nl; k++) {
                    S[(size_t)r * nl + k] -= m * S[(size_t)col * nl + k];
                    I[(size_t)r * nl + k] -= m * I[(size_t)col * nl + k];
                }
            }
        }
        std::vector<float> inv32(I.begin(), I.end());
        HIPCHK(hipMemcpy(c->d_pc_s0inv, inv32.data(), inv32.size() * sizeof(float), hipMemcpyHostToDevice));
        c->pc_ok = 1;
        return 0;
    }
    // mg_precond_level -> mg_relax_vcycle, one cycle from a zero guess
    // (SayakaMacSolver.cpp:245-374, 924-949).  On entry RESID on T(L) already holds the
    // masked right-hand side (rhs - L*0); the zero-guess residual is a masked copy and the
    // residual after the correction is never read (it is only logged), so neither is
    // recomputed here.  phat = 0 + CORR on T(L).
    static int vcycle(sayaka_dev_ctx* c, int L, double* phat) {
        int e;
        double* resid = c->d_field[SAYAKA_RESID];
        double* corr = c->d_field[SAYAKA_CORR];
        for (int l = L; l >= 2; l--)
            if ((e = restrict_(c, l - 1, resid))) return e;
        if ((e = bottom(c, corr, resid, 1.0e-6, 1.0e-6, 2000))) return e;
        for (int l = 2; l <= L; l++) {
            if ((e = prolong(c, l, corr))) return e;
            for (int s = 0; s < 2; s++)
                if ((e = smooth(c, l, corr, resid))) return e;
        }
        VecArgs a = vec_args(c, c->tw[L]);
        a.src = corr;
        a.dst = phat;
        return launch_vec<D, NB>(c, c->tw[L], V_COPYPLUS0, a, FIN_NONE);
    }

    // mg_precond_level (SayakaMacSolver.cpp:924-949): zero-guess V-cycle on tower m.
    static int precond(sayaka_dev_ctx* c, int m, double* sol, const double* rhs) {
        VecArgs a = vec_args(c, c->tw[m]);
        a.src = rhs;
        a.dst = c->d_field[SAYAKA_RESID];
        int e = launch_vec<D, NB>(c, c->tw[m], V_MASKCOPY, a, FIN_NONE);
        if (e) return e;
        return vcycle(c, m, sol);
    }

    // ------------------------------------------------------------- mixed precision
    static int mix(sayaka_dev_ctx* c, ListRef lst, int op, float* dst, const float* src, double* dst64,
                   const double* src64) {
        if (lst.n <= 0) return 0;
        hipLaunchKernelGGL((k_mix<D, NB>), dim3(lst.n), dim3(G::THREADS), 0, c->stream, dlist(c, lst), op,
                           c->d_binfo, c->mixed.vf32, dst, src, dst64, src64);
        HIPCHK(hipGetLastError());
        return 0;
    }
    static StencilArgsT<float> args32(sayaka_dev_ctx* c, ListRef lst, int m) {
        StencilArgsT<float> a{};
        a.list = dlist(c, lst);
        a.nlist = lst.n;
        a.m = m;
        a.binfo = c->d_binfo;
        a.met = c->d_met;
        a.w = c->mixed.w32;
        a.wcf = c->mixed.wcf32;
        a.vf = c->mixed.vf32;
        a.alpha = (float)c->alpha;
        return a;
    }
    static int mixed_prepare(sayaka_dev_ctx* c) {
        MixedState& ms = c->mixed;
        const size_t cells = (size_t)c->cap * c->ncell;
        const size_t faces = (size_t)c->cap * D * c->nface;
        const size_t cf = (size_t)c->cap * c->nf * c->nslab;
        if (ms.cap_cells < cells || ms.nlev < c->maxlevel) {
            mixed_free(ms);
            HIPCHK(hipMalloc(&ms.w32, faces * sizeof(float)));
            HIPCHK(hipMalloc(&ms.wcf32, cf * sizeof(float)));
            HIPCHK(hipMalloc(&ms.vf32, cells * sizeof(float)));
            HIPCHK(hipMalloc(&ms.t32, cells * sizeof(float)));
            const size_t slabs = (size_t)c->cap * 2 * G::NSLAB;
            HIPCHK(hipMalloc(&ms.xst, slabs * sizeof(float)));
            for (int l = 1; l <= c->maxlevel; l++) {
                HIPCHK(hipMalloc(&ms.u[l], cells * sizeof(float)));
                HIPCHK(hipMalloc(&ms.f[l], cells * sizeof(float)));
                HIPCHK(hipMalloc(&ms.xs[l], slabs * sizeof(float)));
            }
            ms.cap_cells = cells;
            ms.cap_faces = faces;
            ms.cap_cf = cf;
            ms.nlev = c->maxlevel;
            ms.coef_version = -1;
        }
        if (ms.coef_version != c->coef_version) {
            const size_t n1 = (size_t)c->nblk * D * c->nface, n2 = (size_t)c->nblk * c->nf * c->nslab,
                         n3 = (size_t)c->nblk * c->ncell;
            hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(256), 0, c->stream, c->d_w, ms.w32, n1);
            hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(256), 0, c->stream, c->d_wcf, ms.wcf32, n2);
            hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(256), 0, c->stream, c->d_vf, ms.vf32, n3);
            HIPCHK(hipGetLastError());
            ms.coef_version = c->coef_version;
        }
        if (c->pc_version != c->coef_version) {
            int e = pc_prepare(c);
            if (e) return e;
            c->pc_version = c->coef_version;
        }
        return 0;
    }
    static int smooth32(sayaka_dev_ctx* c, int m, float* u, const float* f) {
        if (c->mixed.merged_smooth) {
            // both colours over the whole tower, u -> t32 -> u
            ListRef T = c->tw[m];
            if (T.n <= 0) return 0;
            for (int rb = 0; rb < 2; rb++) {
                StencilArgsT<float> a = args32(c, T, m);
                a.x = rb == 0 ? u : c->mixed.t32;
                a.out = rb == 0 ? c->mixed.t32 : u;
                a.rhs = f;
                a.colour = rb;
                ProfScope ps(c, SAYAKA_KCLASS_SMOOTH32, (long long)T.n * G::NCELL,
                             T.n * AlgBytes<D, NB, float>::stencil(ST_SMOOTH), true);
                hipExtLaunchKernelGGL((k_smooth<D, NB, ST_SMOOTH_DB, float>), dim3(T.n), dim3(SmoothCfg<D, NB>::NT), 0,
                                      c->stream, ps.start(), ps.stop(), 0, a);
            }
            HIPCHK(hipGetLastError());
            return 0;
        }
        for (int l = m; l >= 1; l--) {
            ListRef sub = c->tw_sub[m][l];
            if (sub.n <= 0) break;
            for (int rb = 0; rb < 2; rb++) {
                StencilArgsT<float> a = args32(c, sub, m);
                a.x = u;
                a.rhs = f;
                a.colour = rb;
                ProfScope ps(c, SAYAKA_KCLASS_SMOOTH32, (long long)sub.n * G::NCELL,
                             sub.n * AlgBytes<D, NB, float>::stencil(ST_SMOOTH), true);
                hipExtLaunchKernelGGL((k_smooth<D, NB, ST_SMOOTH, float>), dim3(sub.n), dim3(SmoothCfg<D, NB>::NT), 0,
                                      c->stream, ps.start(), ps.stop(), 0, a);
            }
        }
        HIPCHK(hipGetLastError());
        return 0;
    }
    // n hybrid sweeps (k_sweep32) on T(m), alternating between u and t32 so that the last one
    // writes u: the initial value (not read when from_zero) is in sweep_start(u, n).  out64
    // (may be null) receives the result in FP64 too; *wrote64 = 1 when the last sweep wrote it.
    static float* sweep_start(sayaka_dev_ctx* c, float* u, int n) { return (n & 1) ? c->mixed.t32 : u; }
    // rhs64 (from_zero only): the first sweep reads the right-hand side in FP64 and stores f.
    static int sweeps32(sayaka_dev_ctx* c, int m, float* u, float* f, int n, bool from_zero, double* out64,
                        int* wrote64, const double* rhs64 = nullptr) {
        *wrote64 = 0;
        ListRef T = c->tw[m];
        if (T.n <= 0 || n <= 0) return 0;
        float* cur = sweep_start(c, u, n);
        float* nxt = cur == u ? c->mixed.t32 : u;
        // the x-boundary columns of a buffer: t32's, or those of u = u[m]
        auto slab = [&](const float* buf) { return buf == c->mixed.t32 ? c->mixed.xst : c->mixed.xs[m]; };
        for (int s = 0; s < n; s++) {
            const bool zero = s == 0 && from_zero;
            StencilArgsT<float> a = args32(c, T, m);
            a.x = zero ? f : cur;   // not read from zero; any valid array
            a.out = nxt;
            a.rhs = f;
            a.out2 = f;
            const double* r64 = zero ? rhs64 : nullptr;
            double* o64 = s == n - 1 ? out64 : nullptr;
            using AB = AlgBytes<D, NB, float>;
            // a sweep reads x, rhs, vf, w and writes x; from zero x is not read; the FP64
            // right-hand side adds 4 B read and 4 B written per cell
            const double bytes = 2 * AB::stencil(ST_SMOOTH) - (zero ? AB::C : 0.0) + (r64 ? 2 * AB::C : 0.0);
            ProfScope ps(c, SAYAKA_KCLASS_SMOOTH32, (long long)T.n * G::NCELL, T.n * bytes, true);
            constexpr int SNT = SweepCfg<D, NB>::NT;
            const float* xsin = zero ? nullptr : slab(cur);
            float* xsout = slab(nxt);
            if (zero && r64)
                hipExtLaunchKernelGGL((k_sweep32<D, NB, true, true>), dim3(T.n), dim3(SNT), 0, c->stream,
                                      ps.start(), ps.stop(), 0, a, o64, r64, xsin, xsout);
            else if (zero)
                hipExtLaunchKernelGGL((k_sweep32<D, NB, true>), dim3(T.n), dim3(SNT), 0, c->stream, ps.start(),
                                      ps.stop(), 0, a, o64, nullptr, xsin, xsout);
            else
                hipExtLaunchKernelGGL((k_sweep32<D, NB, false>), dim3(T.n), dim3(SNT), 0, c->stream, ps.start(),
                                      ps.stop(), 0, a, o64, nullptr, xsin, xsout);
            if (o64) *wrote64 = 1;
            std::swap(cur, nxt);
        }
        HIPCHK(hipGetLastError());
        return 0;
    }
    // nu sweeps of the configured smoother on T(l) for u_l (from zero when from_zero)
    static int smooth_level(sayaka_dev_ctx* c, int l, int nu, bool from_zero, double* out64, int* wrote64,
                            const double* rhs64 = nullptr) {
        MixedState& ms = c->mixed;
        *wrote64 = 0;
        if (ms.merged_smooth == 2 && nu > 0)
            return sweeps32(c, l, ms.u[l], ms.f[l], nu, from_zero, out64, wrote64, rhs64);
        int e;
        if (from_zero && (e = mix(c, c->tw[l], M_ZERO, ms.u[l], nullptr, nullptr, nullptr))) return e;
        for (int s = 0; s < nu; s++)
            if ((e = smooth32(c, l, ms.u[l], ms.f[l]))) return e;
        return 0;
    }
    // rhs64 (the top level with hybrid pre-sweeps): f_l still in FP64 there, read by the first sweep
    static int vrec(sayaka_dev_ctx* c, int l, int pre, int post, double* out64, const double* rhs64 = nullptr) {
        MixedState& ms = c->mixed;
        int e, wrote64 = 0;
        if (l == 1) {
            double* resid = c->d_field[SAYAKA_RESID];
            double* corr = c->d_field[SAYAKA_CORR];
            if ((e = mix(c, c->tw[1], M_F32_TO_F64, nullptr, ms.f[1], resid, nullptr))) return e;
            if ((e = bottom(c, corr, resid, ms.bottom_tol_rel, ms.bottom_tol_abs, ms.bottom_max_iter, c->pc_ok != 0)))
                return e;
            if ((e = mix(c, c->tw[1], M_F64_TO_F32, ms.u[1], nullptr, nullptr, corr))) return e;
            return out64 ? mix(c, c->tw[1], M_F32_TO_F64, nullptr, ms.u[1], out64, nullptr) : 0;
        }
        ListRef T = c->tw[l];
        if ((e = smooth_level(c, l, pre, true, nullptr, &wrote64, rhs64))) return e;
        if (T.n > 0) {
            // f_{l-1} from r = f_l - A_l u_l: copied on the shared leaves, restricted onto the
            // non-leaf blocks of level l-1 (every child of which is a block of T(l))
            StencilArgsT<float> a = args32(c, T, l);
            a.x = ms.u[l];
            a.rhs = ms.f[l];
            // the slab of u_l is current only after hybrid pre-sweeps
            const float* xsin = ms.merged_smooth == 2 && pre > 0 ? ms.xs[l] : nullptr;
            hipLaunchKernelGGL((k_resid_restrict32<D, NB>), dim3(T.n), dim3(G::THREADS), 0, c->stream, a,
                               ms.f[l - 1], l, xsin);
            HIPCHK(hipGetLastError());
        }
        if ((e = vrec(c, l - 1, pre, post, nullptr))) return e;
        // hybrid sweeps: post-sweeps beyond those asked for make up for the stale block boundaries
        // (measured, mixed solve from zero: 16^3 blocks, 13 steps of the moving cycle: hybrid V(2,2)
        // 149 ms, V(2,3) 134-139, V(2,4) 132-134, V(2,5) 142, exact V(2,2) 170; 8^3 blocks, the 3D
        // benchmark: V(2,3) 5.19 ms, V(2,4) 4.95, V(2,5) 4.75, V(2,6) 4.90, exact V(2,2) 5.80)
        const bool hybrid = ms.merged_smooth == 2 && post > 0;
        const int npost = post + (hybrid ? (NB >= 16 ? 1 : 3) : 0);
        if (T.n > 0) {
            // the corrected u goes where the sweeps start (the other buffer for an odd count)
            float* dst = hybrid ? sweep_start(c, ms.u[l], npost) : ms.u[l];
            float* xsout = hybrid ? (dst == ms.t32 ? ms.xst : ms.xs[l]) : nullptr;
            hipLaunchKernelGGL((k_correct32<D, NB>), dim3(T.n), dim3(G::THREADS), 0, c->stream, dlist(c, T),
                               c->d_binfo, ms.vf32, ms.u[l], ms.u[l - 1], dst, l, xsout);
            HIPCHK(hipGetLastError());
        }
        if ((e = smooth_level(c, l, npost, false, out64, &wrote64))) return e;
        if (out64 && !wrote64) return mix(c, T, M_F32_TO_F64, nullptr, ms.u[l], out64, nullptr);
        return 0;
    }
    static int mixed_vcycle(sayaka_dev_ctx* c, int L, const double* resid64, double* phat64, int pre, int post) {
        MixedState& ms = c->mixed;
        int e;
        // hybrid pre-sweeps read the FP64 residual themselves (and store f_L); otherwise convert
        const bool fused = ms.merged_smooth == 2 && pre > 0 && L >= 2;
        if (!fused && (e = mix(c, c->tw[L], M_F64_TO_F32, ms.f[L], nullptr, nullptr, resid64))) return e;
        return vrec(c, L, pre, post, phat64, fused ? resid64 : nullptr);
    }

    // Coarse/fine synchronisation of one or two face arrays [blk][D][nface] (fd1 may be null)
    // on the finest tower (the coarse side of every coarse/fine face gets the average of the
    // fine faces), then the faces of non-leaf blocks averaged from their children, coarsest
    // last (coef.h).  Directions and arrays are independent, so each launch covers all of them
    // (grid y = array * D + direction); the levels keep their order.
    static void sync_faces(sayaka_dev_ctx* c, double* fd0, double* fd1 = nullptr) {
        const int L = c->finest;
        const dim3 gy(1, fd1 ? 2 * D : D);
        for (int l = L; l >= 2; l--) {
            ListRef fine = c->tw_sub[L][l], crse = c->tw_sub[L][l - 1];
            if (c->tw[L].n == 0 || fine.n <= 0 || crse.n <= 0) continue;
            hipLaunchKernelGGL((k_sync_fine<D, NB>), dim3(fine.n, gy.y), dim3(64), 0, c->stream, dlist(c, fine),
                               c->d_binfo, fd0, fd1);
            hipLaunchKernelGGL((k_sync_crse<D, NB>), dim3(crse.n, gy.y), dim3(64), 0, c->stream, dlist(c, crse),
                               c->d_binfo, fd0, fd1, L);
        }
        for (int l = L - 1; l >= 1; l--) {
            ListRef nl = c->lv_nonleaf[l];
            if (nl.n <= 0) continue;
            hipLaunchKernelGGL((k_restrict_face<D, NB>), dim3(nl.n, gy.y), dim3(256), 0, c->stream, dlist(c, nl),
                               c->d_binfo, fd0, fd1);
        }
    }

    static int set_coefs(sayaka_dev_ctx* c) {
        const int L = c->finest;
        sync_faces(c, c->d_b, c->d_af);
        for (int l = L - 1; l >= 1; l--) {
            ListRef nl = c->lv_nonleaf[l];
            if (nl.n <= 0) continue;
            hipLaunchKernelGGL((k_restrict_cell_avg<D, NB>), dim3(nl.n), dim3(G::THREADS), 0, c->stream,
                               dlist(c, nl), c->d_binfo, c->d_vf);
        }
        hipLaunchKernelGGL((k_weights<D, NB>), dim3(c->all_blocks.n), dim3(256), 0, c->stream,
                           dlist(c, c->all_blocks), c->d_binfo, c->d_met, c->d_b, c->d_af, c->beta, c->d_w,
                           c->d_wcf);
        HIPCHK(hipGetLastError());
        return 0;
    }

    static int fix_bc(sayaka_dev_ctx* c, const double* dphi, int ng, double* rhs) {
        ListRef lst = c->tw[c->finest];
        if (lst.n <= 0) return 0;
        hipLaunchKernelGGL((k_fix_bc<D, NB>), dim3(lst.n), dim3(64), 0, c->stream, dlist(c, lst), c->d_binfo,
                           c->d_met, c->d_b, c->d_af, c->d_vf, c->beta, dphi, ng, rhs);
        HIPCHK(hipGetLastError());
        return 0;
    }

    // The scalars a FIN_RUPDATE finish published (sequence number seq) into h_scal, without
    // waiting for the launches queued after it: spin on the sequence number for 1 ms, then poll
    // it every ~20 us, asking the stream for errors (a failed launch never publishes).
    static int read_published(sayaka_dev_ctx* c, unsigned long long seq) {
        unsigned long long* sq = reinterpret_cast<unsigned long long*>(c->h_mirror + S_NSCALAR);
        const auto t0 = std::chrono::steady_clock::now();
        for (unsigned n = 1;; n++) {
            if (__atomic_load_n(sq, __ATOMIC_ACQUIRE) == seq) break;
            if ((n & 255u) == 0) {
                const auto dt = std::chrono::steady_clock::now() - t0;
                if (dt > std::chrono::milliseconds(1)) {
                    const hipError_t q = hipStreamQuery(c->stream);
                    if (q != hipSuccess && q != hipErrorNotReady) HIPCHK(q);
                    if (q == hipSuccess && __atomic_load_n(sq, __ATOMIC_ACQUIRE) != seq)
                        return fail(-5, "solve: published scalars missing after the stream finished");
                    std::this_thread::sleep_for(std::chrono::microseconds(20));
                }
            }
            __builtin_ia32_pause();
        }
        std::memcpy(c->h_scal, c->h_mirror, S_NSCALAR * sizeof(double));
        return 0;
    }

    static int read_scalars(sayaka_dev_ctx* c) {
        HIPCHK(hipMemcpyAsync(c->h_scal, c->d_scal, S_NSCALAR * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return 0;
    }

    // MacSolver::mg_flux_level (SayakaMacOperator.cpp:161-239) into d_flux (flux.h)
    static int flux(sayaka_dev_ctx* c, int m, int dir, const double* x) {
        if (!c->d_flux) HIPCHK(hipMalloc(&c->d_flux, (size_t)c->cap * c->nface * sizeof(double)));
        HIPCHK(hipMemsetAsync(c->d_flux, 0, (size_t)c->nblk * c->nface * sizeof(double), c->stream));
        ListRef T = c->tw[m];
        if (T.n <= 0) return 0;
        FluxArgs a{dlist(c, T), T.n, m, c->d_binfo, c->d_met, x, dir, c->d_flux};
        ProfScope ps(c, SAYAKA_KCLASS_APPLY, (long long)T.n * G::NCELL, T.n * (AlgBytes<D, NB, double>::C +
                                                                            (double)G::NFACE * sizeof(double)));
        hipLaunchKernelGGL((k_flux<D, NB>), dim3(T.n), dim3(G::THREADS), 0, c->stream, a);
        HIPCHK(hipGetLastError());
        return 0;
    }

    // phi's one-layer face ghosts as mg_flux_level's fill leaves them (flux.h, k_solver_ghosts):
    // tree levels 1..m, a level skipped when the tower's next sub-level is empty
    static int solver_ghosts(sayaka_dev_ctx* c, int m, const double* x, double* out) {
        for (int l = 1; l <= m; l++) {
            if (l < m && c->tw_sub[m][l + 1].n <= 0) continue;
            ListRef lst = c->lv_all[l];
            if (lst.n <= 0) continue;
            hipLaunchKernelGGL((k_solver_ghosts<D, NB>), dim3(lst.n), dim3(256), 0, c->stream, dlist(c, lst),
                               c->d_binfo, x, out);
        }
        HIPCHK(hipGetLastError());
        return 0;
    }

    // regrid data movement of one field (flux.h): gather by origin, then inject the new blocks
    // level by level; idx = [origin (nblk)][new blocks sorted by level]; lv[l] their ranges
    static int remap(sayaka_dev_ctx* c, double* x, int nold, const int* d_origin, const int* d_new,
                     const std::vector<ListRef>& lv) {
        HIPCHK(hipMemcpyAsync(c->d_remap_tmp, x, (size_t)nold * G::NCELL * sizeof(double), hipMemcpyDeviceToDevice,
                              c->stream));
        hipLaunchKernelGGL((k_remap_gather<D, NB>), dim3(c->nblk), dim3(G::THREADS), 0, c->stream, d_origin,
                           c->d_remap_tmp, x);
        for (const ListRef& r : lv)
            if (r.n > 0)
                hipLaunchKernelGGL((k_inject<D, NB>), dim3(r.n), dim3(G::THREADS), 0, c->stream, d_new + r.off,
                                   c->d_binfo, x);
        HIPCHK(hipGetLastError());
        return 0;
    }

    // general ghost fill of a grown cell state (statefill.h): one launch per tree level,
    // coarsest first (FillPatch::fillBoundary, SayakaFillPatch.cpp:17-30)
    static int state_fill(sayaka_dev_ctx* c, FillArgs a, size_t lds_bytes) {
        for (int l = 1; l <= c->maxlevel; l++) {
            const ListRef lst = c->lv_all[l];
            if (lst.n <= 0) continue;
            a.list = dlist(c, lst);
            hipLaunchKernelGGL((k_state_fill<D, NB>), dim3(lst.n), dim3(FILL_NT), lds_bytes, c->stream, a);
        }
        HIPCHK(hipGetLastError());
        return 0;
    }

    // the same fill on face data of one direction (statefill.h, k_face_fill)
    static int face_fill(sayaka_dev_ctx* c, FaceFillArgs a, size_t lds_bytes) {
        for (int l = 1; l <= c->maxlevel; l++) {
            const ListRef lst = c->lv_all[l];
            if (lst.n <= 0) continue;
            a.list = dlist(c, lst);
            hipLaunchKernelGGL((k_face_fill<D, NB>), dim3(lst.n), dim3(FILL_NT), lds_bytes, c->stream, a);
        }
        HIPCHK(hipGetLastError());
        return 0;
    }

    // registered-state remap (statefill.h): move, then per level the parents' in-level fill
    // and the new blocks' prolongation.  lists: per level (1..maxlevel) [parents | new blocks]
    // at offsets po/no with counts pn/nn, all inside d_remap.
    static int state_remap(sayaka_dev_ctx* c, const double* old_state, long long old_cstride, double* state,
                           long long cstride, long long ntile, int ncomp, int ngrow, int fdir,
                           const std::vector<int>& po, const std::vector<int>& pn, const std::vector<int>& no,
                           const std::vector<int>& nn, FillArgs fa, FaceFillArgs ffa, ProlongNewArgs pa,
                           size_t lds_fill, size_t lds_face) {
        hipLaunchKernelGGL(k_state_move, dim3(c->nblk, ncomp), dim3(256), 0, c->stream, c->d_remap, old_state,
                           old_cstride, state, cstride, ntile);
        for (int l = 1; l <= c->maxlevel; l++) {
            if (nn[l] == 0) continue;
            if (pn[l] > 0) {
                if (fdir < 0) {
                    fa.list = c->d_remap + po[l];
                    hipLaunchKernelGGL((k_state_fill<D, NB>), dim3(pn[l]), dim3(FILL_NT), lds_fill, c->stream, fa);
                } else {
                    ffa.list = c->d_remap + po[l];
                    hipLaunchKernelGGL((k_face_fill<D, NB>), dim3(pn[l]), dim3(FILL_NT), lds_fill, c->stream, ffa);
                }
            }
            pa.list = c->d_remap + no[l];
            if (fdir < 0)
                hipLaunchKernelGGL((k_prolong_new<D, NB>), dim3(nn[l]), dim3(FILL_NT), 0, c->stream, pa);
            else
                hipLaunchKernelGGL((k_prolong_new_face<D, NB>), dim3(nn[l]), dim3(FILL_NT), lds_face, c->stream, pa);
        }
        HIPCHK(hipGetLastError());
        return 0;
    }

    // valid cells of a grown state in the output file's order (statefill.h, k_pack_levels):
    // one launch over the blocks in level order
    static int pack_levels(sayaka_dev_ctx* c, const double* state, long long comp_stride, int ncomp, int ngrow,
                           double* out) {
        const ListRef lst = c->by_level;
        if (lst.n <= 0) return 0;
        hipLaunchKernelGGL((k_pack_levels<D, NB>), dim3(lst.n * ncomp), dim3(256), 0, c->stream, dlist(c, lst), ncomp,
                           state, comp_stride, ngrow, out);
        HIPCHK(hipGetLastError());
        return 0;
    }

    // seeded face velocity of every block into d_vel (projection.h); d_lo = [nblk][3] corners
    static int gen_velocity(sayaka_dev_ctx* c, const double* d_lo, const double* dlo, const double* root,
                            unsigned long long seed, int zero_walls) {
        VelArgs a{d_lo, {0, 0, 0}, {1, 1, 1}, seed, zero_walls};
        for (int d = 0; d < 3; d++) {
            a.dlo[d] = dlo[d];
            a.root[d] = root[d];
        }
        hipLaunchKernelGGL((k_face_velocity<D, NB>), dim3(c->nblk), dim3(256), 0, c->stream, a, c->d_binfo, c->d_vel);
        HIPCHK(hipGetLastError());
        return 0;
    }

    // out = div(a u) / (dv vf) on every block minus its volume-weighted mean over the fluid
    // cells of the leaves; d_part = 2 * nblk doubles of scratch
    static int divergence(sayaka_dev_ctx* c, double* out, double* d_part, double* mean_out) {
        const ListRef all = c->all_blocks;
        hipLaunchKernelGGL((k_divergence<D, NB>), dim3(all.n), dim3(G::THREADS), 0, c->stream, dlist(c, all),
                           c->d_binfo, c->d_met, c->d_af, c->d_vel, c->d_vf, out, d_part);
        HIPCHK(hipGetLastError());
        std::vector<double> part(2 * (size_t)all.n);
        HIPCHK(hipMemcpyAsync(part.data(), d_part, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        double s = 0.0, w = 0.0;
        for (int q = 0; q < all.n; q++) {
            s += part[2 * q];
            w += part[2 * q + 1];
        }
        const double mean = w > 0 ? s / w : 0.0;
        hipLaunchKernelGGL(cnn::k_sub_mean, dim3(1024), dim3(256), 0, c->stream, out, c->d_vf, mean, c->nblk * G::NCELL);
        HIPCHK(hipGetLastError());
        if (mean_out) *mean_out = mean;
        return 0;
    }

    // u -= B flux(phi) on the faces of the finest tower, faces re-synchronised; the largest
    // |divergence| over the tower before and after (means removed)
    static int project(sayaka_dev_ctx* c, const double* phi, double* d_part, double* before, double* after) {
        const int L = c->finest;
        const ListRef T = c->tw[L];
        int e;
        if ((e = divergence(c, c->d_div, d_part, nullptr))) return e;
        if ((e = dot(c, L, c->d_div, c->d_div, nullptr, before))) return e;
        for (int d = 0; d < D; d++) {
            if ((e = flux(c, L, d, phi))) return e;
            hipLaunchKernelGGL((k_vel_correct<D, NB>), dim3(T.n), dim3(256), 0, c->stream, dlist(c, T), c->d_b,
                               c->d_flux, c->d_vel, d);
            HIPCHK(hipGetLastError());
        }
        sync_faces(c, c->d_vel);
        if ((e = divergence(c, c->d_div, d_part, nullptr))) return e;
        return dot(c, L, c->d_div, c->d_div, nullptr, after);
    }

    static int dot(sayaka_dev_ctx* c, int m, const double* x, const double* y, double* dotv, double* mx) {
        VecArgs a = vec_args(c, c->tw[m]);
        a.src = x;
        a.t = y;
        int e = launch_vec<D, NB>(c, c->tw[m], V_DOT, a, FIN_DOTNORM);
        if (e) return e;
        if ((e = read_scalars(c))) return e;
        if (dotv) *dotv = c->h_scal[S_DOT];
        if (mx) *mx = c->h_scal[S_NORM];
        return 0;
    }

    // dot over tower m with the partials and scalars in `scr` (the GPU's scratch array), the
    // two results copied to out2 on the device; nothing allocated, nothing waited for
    static int dot_noalloc(sayaka_dev_ctx* c, int m, const double* x, const double* y, double* out2, double* scr) {
        ListRef T = c->tw[m];
        RedBuf rb{scr, scr + 3 * (size_t)NOALLOC_MAX_BLOCKS, c->exact};
        if (T.n > 0) {
            VecArgs a = vec_args(c, T);
            a.src = x;
            a.t = y;
            hipLaunchKernelGGL((k_vec<G::NCELL, G::THREADS>), dim3(T.n), dim3(G::THREADS), 0, c->stream, a, (int)V_DOT,
                               rb, (int)FIN_DOTNORM);
        }
        hipLaunchKernelGGL(k_red_finish, dim3(1), dim3(RED_NT), 0, c->stream, rb, T.n, 1, (int)FIN_DOTNORM);
        hipLaunchKernelGGL(k_copy_dot_norm, dim3(1), dim3(64), 0, c->stream, (const double*)rb.scal, out2);
        HIPCHK(hipGetLastError());
        return 0;
    }

    // MacSolver::cg_solve_level on the finest tower (SayakaMacSolver.cpp:683-920), scalars
    // kept on the device; the host reads them back only at the two convergence tests of
    // each iteration.
    static int solve(sayaka_dev_ctx* c, const sayaka_solve_params* P, sayaka_solve_stats* st) {
        const int L = c->finest;
        ListRef T = c->tw[L];
        double** F = c->d_field;
        int e;
        const bool mixed = P->precision == SAYAKA_MIXED;
        if (mixed && (e = mixed_prepare(c))) return e;
        struct ExactScope {
            sayaka_dev_ctx* c;
            int saved;
            ~ExactScope() { c->exact = saved; }
        } exact_scope{c, c->exact};
        c->exact = P->precision == SAYAKA_FP64 ? 1 : 0;
        int precond = P->precond;
        if (precond && L == 1) precond = 0;
        HIPCHK(hipMemsetAsync(c->d_bot_stats, 0, 8 * sizeof(int), c->stream));
        st->status = 2;
        st->iters = 0;
        st->half_step = 0;
        st->nhist = 0;
        // x0 = SOL ; r = RHS - A x0 ; rhat = r ; x = 0 ; rnorm0 = |r| ; rho = rhat.r
        {
            StencilArgs a = stencil_args(c, T, L);
            a.x = F[SAYAKA_SOL];
            a.rhs = F[SAYAKA_RHS];
            a.out = F[SAYAKA_R];
            a.out2 = F[SAYAKA_RHAT];
            a.out3 = F[SAYAKA_X];
            if ((e = launch_stencil<D, NB>(c, ST_RESID_INIT, a, FIN_INIT))) return e;
        }
        if ((e = read_scalars(c))) return e;
        const double rnorm0 = c->h_scal[S_RNORM0];
        st->rnorm0 = rnorm0;
        st->rnorm = rnorm0;
        if (rnorm0 == 0.0) {
            st->status = 0;
            return 0;
        }
        auto push = [&](double v) {
            if (st->hist && st->nhist < P->hist_cap) st->hist[st->nhist] = v;
            st->nhist++;
        };
        bool conv = false;
        int stat = 0, iter;
        double* phat = precond ? F[SAYAKA_PHAT] : F[SAYAKA_P];
        double* shat = precond ? F[SAYAKA_SHAT] : F[SAYAKA_S];
        // The iterate x - x0 alternates between X and X0: the half step updates it in place, the
        // full step writes the other field, so the half step's x survives the second half.
        //
        // With a preconditioner, and while the last norm the host has tested is at least
        // defer_factor x the stopping threshold (a half-iteration that far from convergence
        // practically never stops), the solve queues a half-iteration before it has read whether
        // the previous one stopped: the second half before it reads |s|, the next iteration's
        // first half before it reads |r|.  The finishes of the two reductions then evaluate the
        // host's stopping tests on the device (RedBuf::stop), the queued V-cycle's coarsest-level
        // solve does nothing when they hold, and the queued half step leaves x as it is; the
        // host reads |r| from the copy the finish publishes (read_published) while the queued
        // work runs.  The host applies the same tests to the same values, so every decision, the
        // history, the coarsest-solve counts and the solution are those of testing first.
        double* xc = F[SAYAKA_X];
        double* xo = F[SAYAKA_X0];
        double* xsol = xc;
        const double thr = std::max(rnorm0 * P->tol_rel, P->tol_abs);
        double last_tested = rnorm0;
        struct DeferScope {
            sayaka_dev_ctx* c;
            ~DeferScope() {
                c->stop_flags = c->publish = false;
                c->skip_ptr = nullptr;
            }
        } defer_scope{c};
        c->stop_flags = precond != 0;
        c->stop_tol_rel = P->tol_rel;
        c->stop_tol_abs = P->tol_abs;
        auto far = [&]() { return precond && c->defer_factor >= 0.0 && last_tested >= c->defer_factor * thr; };
        // the half-step test (SayakaMacSolver.cpp:683-920): breakdown, then |s| against the tolerances
        auto half_step_test = [&]() -> bool {
            if (c->h_scal[S_ALPHA_DEN] == 0.0) {
                stat = 1;
                return true;
            }
            const double snorm = c->h_scal[S_SNORM];
            push(-snorm);
            st->rnorm = snorm;
            last_tested = snorm;
            if (snorm < rnorm0 * P->tol_rel || snorm < P->tol_abs) {
                conv = true;
                st->half_step = 1;
                xsol = xc;
                return true;
            }
            return false;
        };
        // first half of an iteration: p, phat = M^-1 p, v = A phat, alpha, s, x += alpha phat, |s|;
        // skip: the stop flag this half is queued behind (its coarsest solve and x update)
        auto first_half = [&](int it, const int* skip) -> int {
            {
                VecArgs a = vec_args(c, T);
                a.r = F[SAYAKA_R];
                a.p = F[SAYAKA_P];
                a.v = F[SAYAKA_V];
                a.resid = F[SAYAKA_RESID];
                if ((e = launch_vec<D, NB>(c, T, it == 1 ? V_PINIT : V_PUPDATE, a, FIN_NONE))) return e;
            }
            if (precond) {
                c->skip_ptr = skip;
                if (mixed) e = mixed_vcycle(c, L, F[SAYAKA_RESID], phat, P->pre_smooth, P->post_smooth);
                else e = vcycle(c, L, phat);
                c->skip_ptr = nullptr;
                if (e) return e;
            }
            {
                StencilArgs a = stencil_args(c, T, L);
                a.x = phat;
                a.out = F[SAYAKA_V];
                a.dotv = F[SAYAKA_RHAT];
                if ((e = launch_stencil<D, NB>(c, ST_APPLY_DOT, a, FIN_ALPHA))) return e;
            }
            VecArgs a = vec_args(c, T);
            a.r = F[SAYAKA_R];
            a.v = F[SAYAKA_V];
            a.s = F[SAYAKA_S];
            a.x = xc;
            a.phat = phat;
            a.resid = F[SAYAKA_RESID];
            a.xskip = skip;
            return launch_vec<D, NB>(c, T, V_SUPDATE, a, FIN_SNORM);
        };
        // second half: shat = M^-1 s, t = A shat, omega, x (into xo) += omega shat, r, |r|, rho
        auto second_half = [&](const int* skip, bool pub) -> int {
            if (precond) {
                c->skip_ptr = skip;
                if (mixed) e = mixed_vcycle(c, L, F[SAYAKA_RESID], shat, P->pre_smooth, P->post_smooth);
                else e = vcycle(c, L, shat);
                c->skip_ptr = nullptr;
                if (e) return e;
            }
            {
                StencilArgs a = stencil_args(c, T, L);
                a.x = shat;
                a.out = F[SAYAKA_T];
                a.dotv = F[SAYAKA_S];
                if ((e = launch_stencil<D, NB>(c, ST_APPLY_DOT2, a, FIN_OMEGA))) return e;
            }
            VecArgs a = vec_args(c, T);
            a.r = F[SAYAKA_R];
            a.s = F[SAYAKA_S];
            a.t = F[SAYAKA_T];
            a.x = xc;
            a.dst = xo;
            a.src = shat;
            a.rhat = F[SAYAKA_RHAT];
            c->publish = pub;
            e = launch_vec<D, NB>(c, T, V_RUPDATE, a, FIN_RUPDATE);
            c->publish = false;
            return e;
        };
        bool queued = false;   // the first half of iteration `iter` is queued already
        for (iter = 1; iter <= P->max_iter; iter++) {
            st->iters = iter;
            if (!queued) {
                if (c->h_scal[S_RHO] == 0.0) { stat = 1; break; }
                if ((e = first_half(iter, nullptr))) return e;
            }
            queued = false;
            const bool defer_half = far();
            if (!defer_half) {
                if ((e = read_scalars(c))) return e;
                if (half_step_test()) break;
            }
            // the next iteration's first half goes in before |r| is read when still far from the
            // threshold (and the iteration limit leaves one); FP64 modes only: the mixed cycle
            // cuts the residual by 1e2-1e4 per iteration, so the factor does not keep its last
            // full step from paying for a queued first half (~0.45 ms of FP32 V-cycle kernels)
            const bool defer_full = !mixed && iter < P->max_iter && far();
            if ((e = second_half(defer_half ? c->d_stop : nullptr, defer_full))) return e;
            const unsigned long long seq = c->pub_seq;
            if (defer_full) {
                // in place on xo, the x the full step writes (left as it is when the solve stops)
                std::swap(xc, xo);
                e = first_half(iter + 1, c->d_stop + 1);
                std::swap(xc, xo);
                if (e) return e;
                if ((e = read_published(c, seq))) return e;
                queued = true;
            } else if ((e = read_scalars(c))) {
                return e;
            }
            if (defer_half && half_step_test()) break;
            if (c->h_scal[S_TT] == 0.0) { stat = 1; break; }
            const double rnorm = c->h_scal[S_RNORM];
            push(rnorm);
            st->rnorm = rnorm;
            last_tested = rnorm;
            if (rnorm < rnorm0 * P->tol_rel + P->tol_abs) {
                conv = true;
                xsol = xo;
                break;
            }
            if (c->h_scal[S_OMEGA] == 0.0) { stat = 1; break; }
            std::swap(xc, xo);
            if (queued && c->h_scal[S_RHO] == 0.0) {   // the next iteration's opening test
                st->iters = iter + 1;
                stat = 1;
                break;
            }
        }
        if (conv) {
            VecArgs a = vec_args(c, T);
            a.x = xsol;
            a.dst = F[SAYAKA_SOL];
            if ((e = launch_vec<D, NB>(c, T, V_FINAL, a, FIN_NONE))) return e;
        }
        HIPCHK(hipMemcpyAsync(c->h_bot_stats, c->d_bot_stats, 8 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        note_bottom_stats(c);
        if (c->h_bot_stats[4]) return fail(-5, "coarsest-level solve: a grid barrier timed out");
        st->bottom_calls = c->h_bot_stats[0];
        st->bottom_iters = c->h_bot_stats[1];
        st->status = conv ? 0 : (stat ? 1 : 2);
        return st->status;
    }
};

int ensure_coefs(sayaka_dev_ctx* c);

// The (dimension, block size) instance of an entry point that reads no coefficient.
template <class F>
int dispatch_nc(sayaka_dev_ctx* c, F&& f) {
    if (c->ndim == 3 && c->nb == 8) return f(Impl<3, 8>{});
    if (c->ndim == 2 && c->nb == 8) return f(Impl<2, 8>{});
    if (c->ndim == 3 && c->nb == 16) return f(Impl<3, 16>{});
    if (c->ndim == 2 && c->nb == 16) return f(Impl<2, 16>{});
    if (c->ndim == 3 && c->nb == 4) return f(Impl<3, 4>{});
    return fail(-2, "unsupported (ndim, nb)");
}

// ... of an entry point that may read them: unit coefficients left pending are derived first.
template <class F>
int dispatch(sayaka_dev_ctx* c, F&& f) {
    if (c->coef_pending)
        if (int e = ensure_coefs(c)) return e;
    if (c->ndim == 3 && c->nb == 8) return f(Impl<3, 8>{});
    if (c->ndim == 2 && c->nb == 8) return f(Impl<2, 8>{});
    if (c->ndim == 3 && c->nb == 16) return f(Impl<3, 16>{});
    if (c->ndim == 2 && c->nb == 16) return f(Impl<2, 16>{});
    if (c->ndim == 3 && c->nb == 4) return f(Impl<3, 4>{});
    return fail(-2, "unsupported (ndim, nb)");
}

bool field_ok(int f) { return f >= 0 && f < SAYAKA_NFIELD; }

// Unit coefficients (vf = a_f = B = 1) and the weights derived from them for the current
// topology and BCs: what set_topology left pending.
int ensure_coefs(sayaka_dev_ctx* c) {
    c->coef_pending = 0;
    const size_t cells = (size_t)c->nblk * c->ncell;
    const size_t faces = (size_t)c->nblk * c->ndim * c->nface;
    HIPCHK(hipSetDevice(c->device));
    hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, c->stream, c->d_vf, cells, 1.0);
    hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, c->stream, c->d_b, faces, 1.0);
    hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, c->stream, c->d_af, faces, 1.0);
    HIPCHK(hipGetLastError());
    int e = dispatch(c, [&](auto I) { return decltype(I)::set_coefs(c); });
    if (e) return e;
    c->coef_version++;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

}  // namespace

// =============================================================================== C ABI
extern "C" {

const char* sayaka_dev_last_error(void) { return g_err.c_str(); }

int sayaka_dev_create(int device, int ndim, int nb, int max_blocks, sayaka_dev_ctx** out) {
    if (!out) return fail(-1, "null out");
    *out = nullptr;
    if (!((ndim == 3 || ndim == 2) && (nb == 8 || nb == 16)) && !(ndim == 3 && nb == 4))
        return fail(-1, "supported blocks: 8^d, 16^d (d = 2, 3) and 4^3");
    if (max_blocks < 1) return fail(-1, "max_blocks < 1");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(-1, "no such device");
    HIPCHK(hipSetDevice(device));
    auto* c = new sayaka_dev_ctx;
    c->device = device;
    c->ndim = ndim;
    c->nb = nb;
    c->cap = max_blocks;
    const int nbz = ndim == 3 ? nb : 1;
    c->ncell = nb * nb * nbz;
    c->nface = (nb + 1) * nb * nbz;
    c->nslab = nb * nbz;
    c->nf = 2 * ndim;
    hipError_t e = scratch_lookup(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    const size_t cells = (size_t)max_blocks * c->ncell;
    const size_t faces = (size_t)max_blocks * ndim * c->nface;
    for (int f = 0; f < SAYAKA_NFIELD && e == hipSuccess; f++) e = hipMalloc(&c->d_field[f], cells * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&c->d_vf, cells * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&c->d_af, faces * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&c->d_b, faces * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&c->d_w, faces * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&c->d_wcf, (size_t)max_blocks * c->nf * c->nslab * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&c->d_partials, (size_t)(max_blocks + 64) * 3 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&c->d_scal, S_NSCALAR * sizeof(double));
    if (e == hipSuccess) e = hipHostMalloc(&c->h_scal, S_NSCALAR * sizeof(double), hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&c->d_bot_vpub, cells * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&c->d_bot_rpub, cells * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&c->d_bot_part, (size_t)(max_blocks + 64) * 12 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&c->d_l1pos, (size_t)max_blocks * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&c->d_pc_jac, (size_t)PC_MAXBLK * c->ncell * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&c->d_pc_fsum, (size_t)PC_MAXBLK * 6 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&c->d_pc_s0inv, (size_t)PC_MAXBLK * PC_MAXBLK * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&c->d_bot_stats, 8 * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&c->d_stop, 2 * sizeof(int));
    if (e == hipSuccess) e = hipMemset(c->d_stop, 0, 2 * sizeof(int));
    if (e == hipSuccess) e = hipHostMalloc(&c->h_mirror, (S_NSCALAR + 1) * sizeof(double), hipHostMallocCoherent);
    if (e == hipSuccess) {
        std::memset(c->h_mirror, 0, (S_NSCALAR + 1) * sizeof(double));
        e = hipHostGetDevicePointer((void**)&c->d_mirror, c->h_mirror, 0);
    }
    if (e == hipSuccess) e = hipMalloc(&c->d_bot_flags, (size_t)(max_blocks + 64) * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(c->d_bot_flags, 0, (size_t)(max_blocks + 64) * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipHostMalloc(&c->h_bot_stats, 8 * sizeof(int), hipHostMallocDefault);
    if (e == hipSuccess) e = hipMemsetAsync(c->d_scal, 0, S_NSCALAR * sizeof(double), c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->d_wcf, 0, (size_t)max_blocks * c->nf * c->nslab * sizeof(double), c->stream);
    for (int f = 0; f < SAYAKA_NFIELD && e == hipSuccess; f++) e = hipMemsetAsync(c->d_field[f], 0, cells * sizeof(double), c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        std::string msg = std::string("allocation failed: ") + hipGetErrorString(e);
        sayaka_dev_destroy(c);
        return fail(-3, msg);
    }
    *out = c;
    return 0;
}

int sayaka_dev_destroy(sayaka_dev_ctx* c) {
    if (!c) return 0;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto& p : c->d_field) if (p) hipFree(p);
    if (c->d_flux) hipFree(c->d_flux);
    if (c->d_ghost) hipFree(c->d_ghost);
    if (c->d_remap_tmp) hipFree(c->d_remap_tmp);
    if (c->d_remap_idx) hipFree(c->d_remap_idx);
    if (c->d_move) hipFree(c->d_move);
    if (c->d_vel) hipFree(c->d_vel);
    if (c->d_div) hipFree(c->d_div);
    if (c->d_surr) hipFree(c->d_surr);
    if (c->d_bcref) hipFree(c->d_bcref);
    if (c->d_fill_int) hipFree(c->d_fill_int);
    if (c->d_remap) hipFree(c->d_remap);
    if (c->d_fill_val) hipFree(c->d_fill_val);
    void* ptrs[] = {c->d_vf, c->d_af, c->d_b, c->d_w, c->d_wcf, c->d_partials, c->d_scal,
                    c->d_bot_vpub, c->d_bot_rpub, c->d_bot_part, c->d_bot_stats, c->d_stop, c->d_bot_flags, c->d_binfo, c->d_met, c->d_lists,
                    c->d_l1pos, c->d_pc_jac, c->d_pc_fsum, c->d_pc_s0inv};
    for (void* p : ptrs) if (p) hipFree(p);
    for (int k = 0; k < SAYAKA_KCLASS_N; k++)
        for (auto& r : c->prof[k]) {
            hipEventDestroy(r.a);
            hipEventDestroy(r.b);
        }
    for (auto e : c->ev_pool) hipEventDestroy(e);
    if (c->h_scal) hipHostFree(c->h_scal);
    if (c->h_mirror) hipHostFree(c->h_mirror);
    if (c->h_bot_stats) hipHostFree(c->h_bot_stats);
    mixed_free(c->mixed);
    cnn_free(c->cnn);
    if (c->d_gen) (void)hipFree(c->d_gen);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

int sayaka_dev_set_topology(sayaka_dev_ctx* c, int nblk, int minlevel, int maxlevel, const int* level,
                            const int* parent, const int* which_child, const int* child8, const int* nbr6,
                            const int* node_type, const double* length3) {
    if (!c) return fail(-1, "null ctx");
    if (nblk < 1 || nblk > c->cap) return fail(-1, "nblk exceeds the context capacity");
    if (minlevel < 2) return fail(-1, "minAmrLevel must be >= 2 (level-1 leaves alias the outer Krylov state)");
    if (maxlevel < minlevel || maxlevel > MAXLV - 1) return fail(-1, "bad maxlevel");
    HIPCHK(hipSetDevice(c->device));
    c->nblk = nblk;
    c->minlevel = minlevel;
    c->maxlevel = maxlevel;
    c->h_binfo.assign(nblk, BlockInfo{});
    c->h_met.assign(nblk, BlockMetric{});
    int finest = 0;
    for (int b = 0; b < nblk; b++) {
        if (level[b] < 1 || level[b] > maxlevel) return fail(-1, "block level out of range");
        finest = std::max(finest, level[b]);
    }
    c->finest = finest;
    for (int b = 0; b < nblk; b++) {
        BlockInfo& bi = c->h_binfo[b];
        bi.level = level[b];
        bi.leaf = node_type[b] == 1;
        bi.parent = parent[b];
        bi.which_child = which_child[b];
        for (int q = 0; q < 8; q++) bi.child[q] = child8[b * 8 + q];
        for (int f = 0; f < 6; f++)
            for (int q = 0; q < 4; q++) bi.nfc[f][q] = -1;
        for (int f = 0; f < 6; f++) {
            const int n = f < c->nf ? nbr6[b * 6 + f] : PHYS_BNDRY - 1;
            bi.nbr[f] = n;
            bi.cnbr[f] = -1;
            bi.nbr_leaf[f] = 0;
            bi.bct[f] = -1;
            if (n >= nblk) return fail(-1, "neighbour id out of range");
            if (n >= 0) {
                bi.nbr_leaf[f] = node_type[n] == 1;
                if (!bi.nbr_leaf[f] && child8[n * 8] >= 0) {
                    // children of the neighbour that touch face f, by tangential sides (s1, s2)
                    const int d = f >> 1, side = f & 1;
                    const int a1 = d == 0 ? 1 : 0, a2 = d == 2 ? 1 : 2;
                    for (int s2 = 0; s2 < (c->ndim == 3 ? 2 : 1); s2++)
                        for (int s1 = 0; s1 < 2; s1++) {
                            int ch = (1 - side) << d;
                            ch |= s1 << a1;
                            if (c->ndim == 3) ch |= s2 << a2;
                            bi.nfc[f][s1 + 2 * s2] = child8[n * 8 + ch];
                        }
                }
            }
            else if (n > PHYS_BNDRY) {
                if (parent[b] < 0) return fail(-1, "coarse face on a root block");
                const int cn = nbr6[parent[b] * 6 + f];
                if (cn < 0) return fail(-1, "coarse neighbour missing (2:1 balance violated)");
                bi.cnbr[f] = cn;
            }
        }
        // block metric (SayakaTree.h:277-301)
        BlockMetric& mt = c->h_met[b];
        double dx[3], v = 1.0;
        for (int d = 0; d < 3; d++) dx[d] = length3[b * 3 + d];
        for (int d = 0; d < c->ndim; d++) dx[d] /= c->nb;
        for (int d = 0; d < c->ndim; d++) v *= dx[d];
        for (int d = 0; d < 3; d++) {
            mt.dx[d] = dx[d];
            mt.ds[d] = d < c->ndim ? v / dx[d] : 1.0;
        }
        mt.dv = v;
    }
    // block lists
    std::vector<int>& L = c->h_lists;
    L.clear();
    auto add = [&](ListRef& r, auto pred) {
        r.off = (int)L.size();
        for (int b = 0; b < nblk; b++)
            if (pred(b)) L.push_back(b);
        r.n = (int)L.size() - r.off;
    };
    for (int l = 0; l <= MAXLV; l++) c->lv_all[l] = c->lv_leaf[l] = c->lv_nonleaf[l] = c->tw[l] = ListRef{};
    for (int l = 1; l <= maxlevel; l++) {
        add(c->lv_all[l], [&](int b) { return level[b] == l; });
        add(c->lv_leaf[l], [&](int b) { return level[b] == l && node_type[b] == 1; });
        add(c->lv_nonleaf[l], [&](int b) { return level[b] == l && node_type[b] != 1; });
    }
    add(c->all_blocks, [&](int) { return true; });
    c->by_level.off = (int)L.size();
    for (int l = 1; l <= maxlevel; l++)
        for (int b = 0; b < nblk; b++)
            if (level[b] == l) L.push_back(b);
    c->by_level.n = (int)L.size() - c->by_level.off;
    for (int m = 1; m <= maxlevel; m++) {
        for (int l = 0; l <= MAXLV; l++) c->tw_sub[m][l] = ListRef{};
        c->tw[m].off = (int)L.size();
        if (c->lv_all[m].n > 0) {
            for (int l = m; l >= 1; l--) {
                ListRef& s = c->tw_sub[m][l];
                s.off = (int)L.size();
                for (int b = 0; b < nblk; b++)
                    if (level[b] == l && (node_type[b] == 1 || l == m)) L.push_back(b);
                s.n = (int)L.size() - s.off;
            }
        }
        c->tw[m].n = (int)L.size() - c->tw[m].off;
    }
    if (c->d_binfo) hipFree(c->d_binfo);
    if (c->d_met) hipFree(c->d_met);
    if (c->d_lists) hipFree(c->d_lists);
    c->d_binfo = nullptr;
    c->d_met = nullptr;
    c->d_lists = nullptr;
    HIPCHK(hipMalloc(&c->d_binfo, nblk * sizeof(BlockInfo)));
    HIPCHK(hipMalloc(&c->d_met, nblk * sizeof(BlockMetric)));
    HIPCHK(hipMalloc(&c->d_lists, std::max<size_t>(1, L.size()) * sizeof(int)));
    HIPCHK(hipMemcpy(c->d_binfo, c->h_binfo.data(), nblk * sizeof(BlockInfo), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_met, c->h_met.data(), nblk * sizeof(BlockMetric), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_lists, L.data(), L.size() * sizeof(int), hipMemcpyHostToDevice));
    {
        std::vector<int> l1pos(nblk, -1);
        for (int q = 0; q < c->tw[1].n; q++) l1pos[L[c->tw[1].off + q]] = q;
        HIPCHK(hipMemcpy(c->d_l1pos, l1pos.data(), nblk * sizeof(int), hipMemcpyHostToDevice));
        c->pc_version = -1;
    }
    c->vel_valid = 0;  // the face velocity belongs to the previous tree
    c->xfer[2]++;
    // prepareSolver: unit coefficients (SayakaMacCoef.cpp:8-18, SayakaMacIB.cpp:16-32), filled
    // and turned into weights by the first entry point that reads them (ensure_coefs) unless
    // the caller sets its own first
    c->alpha = 0.0;
    c->beta = 1.0;
    c->coef_pending = 1;
    c->coef_version++;
    return 0;
}

int sayaka_dev_set_bc(sayaka_dev_ctx* c, const int* bc6) {
    if (!c || !bc6) return fail(-1, "null argument");
    HIPCHK(hipSetDevice(c->device));
    for (int b = 0; b < c->nblk; b++)
        for (int f = 0; f < c->nf; f++) c->h_binfo[b].bct[f] = c->h_binfo[b].nbr[f] <= PHYS_BNDRY ? bc6[b * 6 + f] : -1;
    HIPCHK(hipMemcpy(c->d_binfo, c->h_binfo.data(), c->nblk * sizeof(BlockInfo), hipMemcpyHostToDevice));
    c->pc_version = -1;
    if (c->coef_pending) {   // the weights are derived later, with these BCs
        c->coef_version++;
        return 0;
    }
    // Neumann / Dirichlet change the wall weights
    int e = dispatch(c, [&](auto I) { return decltype(I)::set_coefs(c); });
    c->coef_version++;
    return e;
}

int sayaka_dev_set_coefs(sayaka_dev_ctx* c, double alpha, double beta, const double* vf, const double* const* af,
                         const double* const* b) {
    if (!c) return fail(-1, "null ctx");
    if (alpha != 0.0) return fail(-1, "alpha must be 0 (the reference asserts it, SayakaMacOperator.cpp:1268)");
    HIPCHK(hipSetDevice(c->device));
    c->coef_pending = 0;   // every coefficient array is written below
    c->alpha = alpha;
    c->beta = beta;
    const size_t cells = (size_t)c->nblk * c->ncell;
    const size_t nfd = (size_t)c->nface;
    const size_t faces = (size_t)c->nblk * c->ndim * nfd;
    // unit coefficients are filled on the device (a regridded tree resets them every step)
    if (vf) HIPCHK(hipMemcpy(c->d_vf, vf, cells * sizeof(double), hipMemcpyHostToDevice));
    else hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, c->stream, c->d_vf, cells, 1.0);
    // host arrays are [dir][block][face]; device arrays are [block][dir][face]
    std::vector<double> tmp;
    for (int which = 0; which < 2; which++) {
        const double* const* src = which == 0 ? b : af;
        bool any = false;
        for (int d = 0; d < c->ndim; d++) any = any || (src && src[d]);
        if (!any) {
            hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, c->stream, which == 0 ? c->d_b : c->d_af, faces, 1.0);
            continue;
        }
        tmp.resize(faces);
        for (int blk = 0; blk < c->nblk; blk++)
            for (int d = 0; d < c->ndim; d++) {
                double* o = tmp.data() + ((size_t)blk * c->ndim + d) * nfd;
                if (src && src[d]) std::memcpy(o, src[d] + (size_t)blk * nfd, nfd * sizeof(double));
                else for (size_t q = 0; q < nfd; q++) o[q] = 1.0;
            }
        HIPCHK(hipMemcpy(which == 0 ? c->d_b : c->d_af, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    int e = dispatch(c, [&](auto I) { return decltype(I)::set_coefs(c); });
    if (e) return e;
    c->coef_version++;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int sayaka_dev_get_coefs(sayaka_dev_ctx* c, double* vf, double* const* af, double* const* b, double* w, double* wcf) {
    if (!c) return fail(-1, "null ctx");
    if (c->coef_pending)
        if (int e = ensure_coefs(c)) return e;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    const size_t cells = (size_t)c->nblk * c->ncell;
    const size_t nfd = (size_t)c->nface;
    if (vf) HIPCHK(hipMemcpy(vf, c->d_vf, cells * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<double> tmp((size_t)c->nblk * c->ndim * nfd);
    for (int which = 0; which < 2; which++) {
        double* const* dst = which == 0 ? b : af;
        if (!dst) continue;
        HIPCHK(hipMemcpy(tmp.data(), which == 0 ? c->d_b : c->d_af, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int blk = 0; blk < c->nblk; blk++)
            for (int d = 0; d < c->ndim; d++)
                if (dst[d]) std::memcpy(dst[d] + (size_t)blk * nfd, tmp.data() + ((size_t)blk * c->ndim + d) * nfd, nfd * sizeof(double));
    }
    if (w) HIPCHK(hipMemcpy(w, c->d_w, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
    if (wcf) HIPCHK(hipMemcpy(wcf, c->d_wcf, (size_t)c->nblk * c->nf * c->nslab * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int sayaka_dev_upload(sayaka_dev_ctx* c, int field, const double* h) {
    if (!c || !h || !field_ok(field)) return fail(-1, "bad argument");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(c->d_field[field], h, (size_t)c->nblk * c->ncell * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->xfer[0] += (long long)c->nblk * c->ncell * (long long)sizeof(double);
    return 0;
}

int sayaka_dev_download(sayaka_dev_ctx* c, int field, double* h) {
    if (!c || !h || !field_ok(field)) return fail(-1, "bad argument");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(h, c->d_field[field], (size_t)c->nblk * c->ncell * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->xfer[1] += (long long)c->nblk * c->ncell * (long long)sizeof(double);
    return 0;
}

int sayaka_dev_fix_bc(sayaka_dev_ctx* c, const double* phi, int ng, int rhs_field) {
    if (!c || !phi || ng < 1 || !field_ok(rhs_field)) return fail(-1, "bad argument");
    HIPCHK(hipSetDevice(c->device));
    const int nbz = c->ndim == 3 ? c->nb : 1;
    const size_t n = (size_t)c->nblk * (c->nb + 2 * ng) * (c->nb + 2 * ng) * (nbz + (c->ndim == 3 ? 2 * ng : 0));
    double* d = nullptr;
    HIPCHK(hipMalloc(&d, n * sizeof(double)));
    hipError_t e = hipMemcpy(d, phi, n * sizeof(double), hipMemcpyHostToDevice);
    int r = e == hipSuccess ? dispatch(c, [&](auto I) { return decltype(I)::fix_bc(c, d, ng, c->d_field[rhs_field]); })
                            : fail(-3, hipGetErrorString(e));
    hipStreamSynchronize(c->stream);
    hipFree(d);
    return r;
}

int sayaka_dev_apply(sayaka_dev_ctx* c, int m, int out, int in) {
    if (!c || !field_ok(out) || !field_ok(in) || m < 1 || m > c->maxlevel) return fail(-1, "bad argument");
    int e = dispatch(c, [&](auto I) { return decltype(I)::apply(c, m, out, in); });
    if (e) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int sayaka_dev_residual(sayaka_dev_ctx* c, int m, int r, int x, int rhs) {
    if (!c || !field_ok(r) || !field_ok(x) || !field_ok(rhs) || m < 1 || m > c->maxlevel) return fail(-1, "bad argument");
    int e = dispatch(c, [&](auto I) { return decltype(I)::residual(c, m, r, x, rhs); });
    if (e) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int sayaka_dev_smooth(sayaka_dev_ctx* c, int m, int sol, int rhs) {
    if (!c || !field_ok(sol) || !field_ok(rhs) || m < 1 || m > c->maxlevel) return fail(-1, "bad argument");
    int e = dispatch(c, [&](auto I) { return decltype(I)::smooth(c, m, c->d_field[sol], c->d_field[rhs]); });
    if (e) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int sayaka_dev_restrict(sayaka_dev_ctx* c, int cl, int field) {
    if (!c || !field_ok(field)) return fail(-1, "bad argument");
    int e = dispatch(c, [&](auto I) { return decltype(I)::restrict_(c, cl, c->d_field[field]); });
    if (e) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int sayaka_dev_prolong(sayaka_dev_ctx* c, int fl, int field) {
    if (!c || !field_ok(field)) return fail(-1, "bad argument");
    int e = dispatch(c, [&](auto I) { return decltype(I)::prolong(c, fl, c->d_field[field]); });
    if (e) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int sayaka_dev_alloc(sayaka_dev_ctx* c, size_t bytes, void** out) {
    if (!c || !out || bytes == 0) return fail(-1, "bad argument");
    *out = nullptr;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMalloc(out, bytes));
    return 0;
}

int sayaka_dev_free(sayaka_dev_ctx* c, void* p) {
    if (!c) return fail(-1, "bad argument");
    if (!p) return 0;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipFree(p));
    return 0;
}

int sayaka_dev_memcpy(sayaka_dev_ctx* c, void* dst, const void* src, size_t bytes, int to_device) {
    if (!c || !dst || !src) return fail(-1, "bad argument");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(dst, src, bytes, to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->xfer[to_device ? 0 : 1] += (long long)bytes;
    return 0;
}

int sayaka_dev_set_half_step_defer(sayaka_dev_ctx* c, double factor) {
    if (!c || factor != factor) return fail(-1, "bad argument");
    c->defer_factor = factor;
    return 0;
}

int sayaka_dev_set_bottom_placement(sayaka_dev_ctx* c, int one_xcd) {
    if (!c || (one_xcd != 0 && one_xcd != 1)) return fail(-1, "bad argument");
    c->bot_one_xcd = one_xcd;
    return 0;
}

int sayaka_dev_bottom_placement(sayaka_dev_ctx* c, long long* out3) {
    if (!c || !out3) return fail(-1, "bad argument");
    for (int q = 0; q < 3; q++) out3[q] = c->bot_place[q];
    return 0;
}

int sayaka_dev_transfer_stats(sayaka_dev_ctx* c, long long* out4) {
    if (!c || !out4) return fail(-1, "bad argument");
    for (int q = 0; q < 4; q++) out4[q] = c->xfer[q];
    return 0;
}

int sayaka_dev_block_move(sayaka_dev_ctx* c, const double* src, double* dst, int ncomp, long long comp_stride,
                          long long block_len, int n, int nslot, const int* origin) {
    if (!c || !src || !dst || !origin || ncomp < 1 || block_len < 1 || n < 0 || nslot < n || nslot < 1 ||
        comp_stride < (long long)nslot * block_len || ncomp > 65535)
        return fail(-1, "bad argument");
    if (src == dst) return fail(-1, "the source and the destination must be distinct buffers");
    for (int b = 0; b < n; b++)
        if (origin[b] < -1 || origin[b] >= nslot) return fail(-1, "origin out of range");
    HIPCHK(hipSetDevice(c->device));
    if ((size_t)nslot > c->move_cap) {
        if (c->d_move) HIPCHK(hipFree(c->d_move));
        c->d_move = nullptr;
        HIPCHK(hipMalloc(&c->d_move, (size_t)nslot * sizeof(int)));
        c->move_cap = nslot;
    }
    if (n > 0) HIPCHK(hipMemcpyAsync(c->d_move, origin, (size_t)n * sizeof(int), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_block_move, dim3(nslot, ncomp), dim3(256), 0, c->stream, src, dst, c->d_move, comp_stride,
                       block_len, n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));   // origin is the caller's host array
    return 0;
}

int sayaka_dev_set_surroundings(sayaka_dev_ctx* c, int nblk, const int* surr27, const int* bcref27) {
    if (!c || !surr27 || !bcref27) return fail(-1, "bad argument");
    if (nblk != c->nblk) return fail(-1, "block count differs from the topology");
    for (int e = 0; e < nblk * 27; e++) {
        const int v = surr27[e];
        if (v >= nblk || bcref27[e] < -1 || bcref27[e] >= 27) return fail(-1, "surrounding table out of range");
    }
    HIPCHK(hipSetDevice(c->device));
    if (!c->d_surr) {
        HIPCHK(hipMalloc(&c->d_surr, (size_t)c->cap * 27 * sizeof(int)));
        HIPCHK(hipMalloc(&c->d_bcref, (size_t)c->cap * 27 * sizeof(int)));
    }
    c->h_surr.assign(surr27, surr27 + (size_t)nblk * 27);
    HIPCHK(hipMemcpyAsync(c->d_surr, surr27, (size_t)nblk * 27 * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_bcref, bcref27, (size_t)nblk * 27 * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->surr_n = nblk;
    c->xfer[3]++;
    return 0;
}

// BC register of a cell-state fill (FillArgs.physcode / bctype / bcval): every (surrounding,
// physical code) the tree holds needs a BC for every filled component (the reference exits on
// an unknown mapping, SayakaBoundaryPatch.h:101-110); the tables go to the device.
static int setup_fill_bc(sayaka_dev_ctx* c, int nc, int scomp, int nphys, const int* physbc, const int* mathbc,
                         const double* bcval) {
    std::vector<char> seen((size_t)27 * (nphys + 1), 0);
    for (int b = 0; b < c->nblk; b++)
        for (int q = 0; q < 27; q++) {
            const int v = c->h_surr[(size_t)b * 27 + q];
            if (v > -20) continue;
            int h = -1;
            for (int t = 0; t < nphys; t++)
                if (physbc[t] == v) h = t;
            if (h < 0) return fail(-2, "no BC mapping for physical code " + std::to_string(v));
            if (seen[(size_t)q * (nphys + 1) + h]) continue;
            seen[(size_t)q * (nphys + 1) + h] = 1;
            for (int cc = 0; cc < nc; cc++) {
                const int t = mathbc[((size_t)cc * 27 + q) * nphys + h];
                if (t < 0 || t > 2)
                    return fail(-2, "no BC mapping for physical code " + std::to_string(v) + " at surrounding " +
                                        std::to_string(q) + " of component " + std::to_string(scomp + cc));
            }
        }
    HIPCHK(hipSetDevice(c->device));
    const size_t ni = (size_t)nphys + (size_t)nc * 27 * nphys, nv = (size_t)nc * 27 * nphys;
    if (ni > c->fill_icap) {
        if (c->d_fill_int) HIPCHK(hipFree(c->d_fill_int));
        HIPCHK(hipMalloc(&c->d_fill_int, (ni + 1) * sizeof(int)));
        c->fill_icap = ni;
    }
    if (nv > c->fill_vcap) {
        if (c->d_fill_val) HIPCHK(hipFree(c->d_fill_val));
        HIPCHK(hipMalloc(&c->d_fill_val, (nv + 1) * sizeof(double)));
        c->fill_vcap = nv;
    }
    if (nphys > 0) {
        std::vector<int> hi(ni);
        std::copy(physbc, physbc + nphys, hi.begin());
        std::copy(mathbc, mathbc + nv, hi.begin() + nphys);
        HIPCHK(hipMemcpyAsync(c->d_fill_int, hi.data(), ni * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->d_fill_val, bcval, nv * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));   // the host vector goes out of scope
    }
    return 0;
}

int sayaka_dev_state_fill(sayaka_dev_ctx* c, double* state, int ncomp, int blocks_alloc, int ngrow, int scomp, int nc,
                          int nlayer, const int* prolong, int nphys, const int* physbc, const int* mathbc,
                          const double* bcval) {
    if (!c || !state || ncomp < 1 || ngrow < 1 || scomp < 0 || nc < 0 || scomp + nc > ncomp)
        return fail(-1, "bad argument");
    if (nlayer <= 0 || nc == 0) return 0;   // the reference returns at once for nlayer <= 0
    if (nlayer > ngrow) return fail(-1, "nlayer > ngrow");
    if (nc > FILL_MAXCOMP) return fail(-1, "more than 16 components in one fill");
    if (c->surr_n != c->nblk || c->nblk == 0) return fail(-1, "surrounding tables not set for this topology");
    if (blocks_alloc < c->nblk) return fail(-1, "blocks_alloc below the tree's block count");
    if (nphys < 0 || (nphys > 0 && (!physbc || !mathbc || !bcval)) || !prolong) return fail(-1, "bad argument");
    const int nz = c->ndim == 3 ? c->nb : 1, gz = c->ndim == 3 ? ngrow : 0;
    const long long gx = c->nb + 2 * ngrow;
    const long long ntile = gx * gx * (nz + 2 * gz);
    const size_t lds = (size_t)ntile * sizeof(double);
    if (lds > 64 * 1024) return fail(-1, "grown block above 64 KiB (ngrow too large for this block size)");
    for (int q = 0; q < nc; q++) {
        if (prolong[q] < 0 || prolong[q] > 2) return fail(-1, "unknown prolongation method");
        // centred slopes read one coarse cell beyond the mapped one (the reference exits,
        // SayakaInterpPatch.cpp:233-236)
        if (prolong[q] > 0 && ngrow < 2) return fail(-1, "centred-linear prolongation needs ngrow >= 2");
    }
    if (int e = setup_fill_bc(c, nc, scomp, nphys, physbc, mathbc, bcval)) return e;
    FillArgs a{};
    a.binfo = c->d_binfo;
    a.surr = c->d_surr;
    a.bcref = c->d_bcref;
    a.physcode = c->d_fill_int;
    a.bctype = c->d_fill_int + nphys;
    a.bcval = c->d_fill_val;
    a.state = state;
    a.comp_stride = (long long)blocks_alloc * ntile;
    a.ngrow = ngrow;
    a.nlayer = nlayer;
    a.scomp = scomp;
    a.nc = nc;
    a.nphys = nphys;
    for (int q = 0; q < nc; q++) a.prolong[q] = prolong[q];
    int e = dispatch_nc(c, [&](auto I) { return decltype(I)::state_fill(c, a, lds); });
    if (e) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int sayaka_dev_face_fill(sayaka_dev_ctx* c, double* state, int ncomp, int blocks_alloc, int fdir, int ngrow,
                         int scomp, int nc, int nlayer, const int* prolong) {
    if (!c || !state || ncomp < 1 || ngrow < 1 || scomp < 0 || nc < 0 || scomp + nc > ncomp || !prolong)
        return fail(-1, "bad argument");
    if (fdir < 0 || fdir >= c->ndim) return fail(-1, "face direction beyond the tree's dimension");
    if (nlayer <= 0 || nc == 0) return 0;   // the reference returns at once for nlayer <= 0
    if (nlayer > ngrow) return fail(-1, "nlayer > ngrow");
    if (nc > FILL_MAXCOMP) return fail(-1, "more than 16 components in one fill");
    if (c->surr_n != c->nblk || c->nblk == 0) return fail(-1, "surrounding tables not set for this topology");
    if (blocks_alloc < c->nblk) return fail(-1, "blocks_alloc below the tree's block count");
    const int nz = c->ndim == 3 ? c->nb : 1, gz = c->ndim == 3 ? ngrow : 0;
    const long long e0 = c->nb + 2 * ngrow + (fdir == 0), e1 = c->nb + 2 * ngrow + (fdir == 1);
    const long long ntile = e0 * e1 * (nz + 2 * gz + (fdir == 2));
    const size_t lds = (size_t)ntile * sizeof(double);
    if (lds > 64 * 1024) return fail(-1, "grown face box above 64 KiB (ngrow too large for this block size)");
    for (int q = 0; q < nc; q++) {
        if (prolong[q] < 0 || prolong[q] > 2) return fail(-1, "unknown prolongation method");
        // tangential slopes read one coarse face beyond the mapped one; the mid-face average of
        // an odd outermost layer reads one face beyond the grown box
        if (prolong[q] > 0 && ngrow < 2) return fail(-1, "centred-linear prolongation needs ngrow >= 2");
        if (prolong[q] > 0 && (nlayer & 1) && nlayer == ngrow)
            return fail(-1, "centred-linear face prolongation of an odd nlayer needs nlayer < ngrow");
    }
    // the reference fills no face data next to a wall (SayakaFillPatch.cpp:318-324)
    for (int b = 0; b < c->nblk; b++)
        for (int q = 0; q < 27; q++)
            if (c->h_surr[(size_t)b * 27 + q] <= -20)
                return fail(-2, "face data fill on a tree with physical boundaries (periodic trees only)");
    HIPCHK(hipSetDevice(c->device));
    FaceFillArgs a{};
    a.binfo = c->d_binfo;
    a.surr = c->d_surr;
    a.state = state;
    a.comp_stride = (long long)blocks_alloc * ntile;
    a.fd = fdir;
    a.ngrow = ngrow;
    a.nlayer = nlayer;
    a.scomp = scomp;
    a.nc = nc;
    for (int q = 0; q < nc; q++) a.prolong[q] = prolong[q];
    int e = dispatch_nc(c, [&](auto I) { return decltype(I)::face_fill(c, a, lds); });
    if (e) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int sayaka_dev_state_remap(sayaka_dev_ctx* c, const double* old_state, int old_blocks_alloc, int nold, double* state,
                           int ncomp, int blocks_alloc, int ngrow, int fdir, const int* origin, const int* prolong,
                           int nphys, const int* physbc, const int* mathbc, const double* bcval) {
    if (!c || !old_state || !state || !origin || !prolong || ncomp < 1 || ngrow < 1 || nold < 1)
        return fail(-1, "bad argument");
    if (old_state == state) return fail(-1, "the old and the new state must be distinct buffers");
    if (ncomp > FILL_MAXCOMP) return fail(-1, "more than 16 components");
    if (fdir < -1 || fdir >= c->ndim) return fail(-1, "face direction beyond the tree's dimension");
    if (c->surr_n != c->nblk || c->nblk == 0) return fail(-1, "surrounding tables not set for this topology");
    if (blocks_alloc < c->nblk || old_blocks_alloc < nold) return fail(-1, "block allocation below the block count");
    const int n = c->nblk, nb = c->nb, nd = c->ndim;
    const int nz = nd == 3 ? nb : 1, gz = nd == 3 ? ngrow : 0;
    const long long ntile = (long long)(nb + 2 * ngrow + (fdir == 0)) * (nb + 2 * ngrow + (fdir == 1)) *
                            (nz + 2 * gz + (fdir == 2));
    const size_t lds_fill = (size_t)ntile * sizeof(double);
    if (lds_fill > 64 * 1024) return fail(-1, "grown block above 64 KiB (ngrow too large for this block size)");
    const size_t lds_face = (size_t)(nb + (fdir == 0)) * (nb + (fdir == 1)) * (nz + (fdir == 2)) * sizeof(double);
    for (int q = 0; q < ncomp; q++)
        if (prolong[q] < 0 || prolong[q] > 2) return fail(-1, "unknown prolongation method");
    // per level: the parents of the level's new blocks (ascending) and the new blocks
    std::vector<std::vector<int>> par(c->maxlevel + 1), nw(c->maxlevel + 1);
    std::vector<char> mark(n, 0);
    for (int b = 0; b < n; b++) {
        if (origin[b] >= nold || origin[b] < -1) return fail(-1, "origin out of range");
        if (origin[b] >= 0) continue;
        const BlockInfo& bi = c->h_binfo[b];
        if (bi.parent < 0 || bi.level < 2) return fail(-1, "a new block without a parent");
        nw[bi.level].push_back(b);
        if (!mark[bi.parent]) {
            mark[bi.parent] = 1;
            par[bi.level].push_back(bi.parent);
        }
    }
    for (auto& v : par) std::sort(v.begin(), v.end());
    if (fdir >= 0) {
        // the reference fills no face data next to a wall (SayakaFillPatch.cpp:318-324)
        for (int b = 0; b < n; b++)
            for (int q = 0; q < 27; q++)
                if (c->h_surr[(size_t)b * 27 + q] <= -20)
                    return fail(-2, "face data remap on a tree with physical boundaries (periodic trees only)");
    } else if (int e = setup_fill_bc(c, ncomp, 0, nphys, physbc, mathbc, bcval)) {
        return e;
    }
    std::vector<int> L(origin, origin + n);
    std::vector<int> po(c->maxlevel + 1, 0), pn(c->maxlevel + 1, 0), no(c->maxlevel + 1, 0), nn(c->maxlevel + 1, 0);
    for (int l = 1; l <= c->maxlevel; l++) {
        po[l] = (int)L.size();
        pn[l] = (int)par[l].size();
        L.insert(L.end(), par[l].begin(), par[l].end());
        no[l] = (int)L.size();
        nn[l] = (int)nw[l].size();
        L.insert(L.end(), nw[l].begin(), nw[l].end());
    }
    HIPCHK(hipSetDevice(c->device));
    if (L.size() > c->remap_cap) {
        if (c->d_remap) HIPCHK(hipFree(c->d_remap));
        HIPCHK(hipMalloc(&c->d_remap, L.size() * sizeof(int)));
        c->remap_cap = L.size();
    }
    HIPCHK(hipMemcpyAsync(c->d_remap, L.data(), L.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    const long long cstride = (long long)blocks_alloc * ntile;
    FillArgs fa{};
    fa.binfo = c->d_binfo;
    fa.surr = c->d_surr;
    fa.bcref = c->d_bcref;
    fa.physcode = c->d_fill_int;
    fa.bctype = c->d_fill_int + nphys;
    fa.bcval = c->d_fill_val;
    fa.state = state;
    fa.comp_stride = cstride;
    fa.ngrow = ngrow;
    fa.nlayer = ngrow;
    fa.scomp = 0;
    fa.nc = ncomp;
    fa.nphys = nphys;
    fa.same_level_only = 1;
    FaceFillArgs ffa{};
    ffa.binfo = c->d_binfo;
    ffa.surr = c->d_surr;
    ffa.state = state;
    ffa.comp_stride = cstride;
    ffa.fd = fdir;
    ffa.ngrow = ngrow;
    ffa.nlayer = ngrow;
    ffa.scomp = 0;
    ffa.nc = ncomp;
    ffa.same_level_only = 1;
    ProlongNewArgs pa{};
    pa.origin = c->d_remap;
    pa.binfo = c->d_binfo;
    pa.state = state;
    pa.comp_stride = cstride;
    pa.ngrow = ngrow;
    pa.nc = ncomp;
    pa.fd = fdir;
    for (int q = 0; q < ncomp; q++) fa.prolong[q] = ffa.prolong[q] = pa.prolong[q] = prolong[q];
    int e = dispatch_nc(c, [&](auto I) {
        return decltype(I)::state_remap(c, old_state, (long long)old_blocks_alloc * ntile, state, cstride, ntile,
                                        ncomp, ngrow, fdir, po, pn, no, nn, fa, ffa, pa, lds_fill, lds_face);
    });
    if (e) return e;
    HIPCHK(hipStreamSynchronize(c->stream));   // L goes out of scope
    return 0;
}

int sayaka_dev_pack_levels(sayaka_dev_ctx* c, const double* state, int ncomp, int blocks_alloc, int ngrow,
                           double* out) {
    if (!c || !state || !out || ncomp < 1 || ngrow < 0) return fail(-1, "bad argument");
    if (c->nblk == 0) return fail(-1, "no topology");
    if (blocks_alloc < c->nblk) return fail(-1, "blocks_alloc below the tree's block count");
    const long long gx = c->nb + 2 * ngrow;
    const long long ntile = gx * gx * (c->ndim == 3 ? c->nb + 2 * ngrow : 1);
    HIPCHK(hipSetDevice(c->device));
    int e = dispatch_nc(c, [&](auto I) {
        return decltype(I)::pack_levels(c, state, (long long)blocks_alloc * ntile, ncomp, ngrow, out);
    });
    if (e) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int sayaka_dev_remap(sayaka_dev_ctx* c, const int* origin, int nold, const int* fields, int nfields) {
    if (!c || !origin || !fields || nfields < 0 || nold < 1 || nold > c->cap) return fail(-1, "bad argument");
    for (int q = 0; q < nfields; q++)
        if (!field_ok(fields[q])) return fail(-1, "bad field");
    const int n = c->nblk;
    std::vector<int> idx(origin, origin + n);
    std::vector<ListRef> lv;
    for (int l = 1; l <= c->maxlevel; l++) {
        ListRef r{(int)idx.size(), 0};
        for (int b = 0; b < n; b++) {
            if (origin[b] >= nold) return fail(-1, "origin beyond the old block count");
            if (origin[b] < 0 && c->h_binfo[b].level == l) {
                if (c->h_binfo[b].parent < 0) return fail(-1, "a new block without a parent");
                idx.push_back(b);
            }
        }
        r.n = (int)idx.size() - r.off;
        lv.push_back(r);
    }
    HIPCHK(hipSetDevice(c->device));
    if (!c->d_remap_tmp) HIPCHK(hipMalloc(&c->d_remap_tmp, (size_t)c->cap * c->ncell * sizeof(double)));
    if (!c->d_remap_idx) HIPCHK(hipMalloc(&c->d_remap_idx, (size_t)2 * c->cap * sizeof(int)));
    HIPCHK(hipMemcpyAsync(c->d_remap_idx, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    for (int q = 0; q < nfields; q++) {
        int e = dispatch_nc(c, [&](auto I) {
            return decltype(I)::remap(c, c->d_field[fields[q]], nold, c->d_remap_idx, c->d_remap_idx, lv);
        });
        if (e) return e;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int sayaka_dev_flux(sayaka_dev_ctx* c, int m, int dir, int phi, double* flux_out) {
    if (!c || !field_ok(phi) || m < 1 || m > c->maxlevel || dir < 0 || dir >= c->ndim || !flux_out)
        return fail(-1, "bad argument");
    HIPCHK(hipSetDevice(c->device));
    int e = dispatch(c, [&](auto I) { return decltype(I)::flux(c, m, dir, c->d_field[phi]); });
    if (e) return e;
    HIPCHK(hipMemcpyAsync(flux_out, c->d_flux, (size_t)c->nblk * c->nface * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int sayaka_dev_solver_ghosts(sayaka_dev_ctx* c, int m, int field, double* ghost_out) {
    if (!c || !field_ok(field) || m < 1 || m > c->maxlevel || !ghost_out) return fail(-1, "bad argument");
    HIPCHK(hipSetDevice(c->device));
    const size_t n = (size_t)c->nblk * 2 * c->ndim * c->nslab;
    if (!c->d_ghost) HIPCHK(hipMalloc(&c->d_ghost, (size_t)c->cap * 2 * c->ndim * c->nslab * sizeof(double)));
    HIPCHK(hipMemsetAsync(c->d_ghost, 0, n * sizeof(double), c->stream));
    int e = dispatch_nc(c, [&](auto I) { return decltype(I)::solver_ghosts(c, m, c->d_field[field], c->d_ghost); });
    if (e) return e;
    HIPCHK(hipMemcpyAsync(ghost_out, c->d_ghost, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int sayaka_dev_dot(sayaka_dev_ctx* c, int m, int a, int b, double* dotv, double* mx) {
    if (!c || !field_ok(a) || !field_ok(b) || m < 1 || m > c->maxlevel) return fail(-1, "bad argument");
    return dispatch(c, [&](auto I) { return decltype(I)::dot(c, m, c->d_field[a], c->d_field[b], dotv, mx); });
}

int sayaka_dev_dot_noalloc(sayaka_dev_ctx* c, int m, const double* a, const double* b, double* out2) {
    if (!c || !a || !b || !out2 || m < 1 || m > c->maxlevel) return fail(-1, "bad argument");
    if (c->tw[m].n > NOALLOC_MAX_BLOCKS)
        return fail(-1, "sayaka_dev_dot_noalloc: tower of " + std::to_string(c->tw[m].n) + " blocks exceeds the " +
                            std::to_string(NOALLOC_MAX_BLOCKS) + "-block scratch");
    double* scr = (c->device >= 0 && c->device < SCRATCH_MAX_DEV) ? scratch_dev[c->device] : nullptr;
    if (!scr) return fail(-3, "sayaka_dev_dot_noalloc: no scratch address for this device");
    HIPCHK(hipSetDevice(c->device));
    return dispatch(c, [&](auto I) { return decltype(I)::dot_noalloc(c, m, a, b, out2, scr); });
}

int sayaka_dev_precond(sayaka_dev_ctx* c, int m, int sol, int rhs) {
    if (!c || !field_ok(sol) || !field_ok(rhs) || m < 2 || m > c->finest) return fail(-1, "bad argument");
    int e = dispatch(c, [&](auto I) { return decltype(I)::precond(c, m, c->d_field[sol], c->d_field[rhs]); });
    if (e) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int sayaka_dev_bottom(sayaka_dev_ctx* c, int sol, int rhs, double tol_rel, double tol_abs, int max_iter, int* iters,
                      int* status) {
    if (!c || !field_ok(sol) || !field_ok(rhs) || max_iter < 1) return fail(-1, "bad argument");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemsetAsync(c->d_bot_stats, 0, 8 * sizeof(int), c->stream));
    int e = dispatch(c, [&](auto I) {
        return decltype(I)::bottom(c, c->d_field[sol], c->d_field[rhs], tol_rel, tol_abs, max_iter);
    });
    if (e) return e;
    HIPCHK(hipMemcpyAsync(c->h_bot_stats, c->d_bot_stats, 8 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    note_bottom_stats(c);
    if (c->h_bot_stats[4]) return fail(-5, "coarsest-level solve: a grid barrier timed out");
    if (iters) *iters = c->h_bot_stats[2];
    if (status) *status = c->h_bot_stats[3];
    return 0;
}

int sayaka_dev_solve(sayaka_dev_ctx* c, const sayaka_solve_params* p, sayaka_solve_stats* st) {
    if (!c || !p || !st) return fail(-1, "null argument");
    if (p->max_iter < 1 || p->tol_rel < 0 || p->tol_abs < 0) return fail(-1, "bad solver parameters");
    if (p->precision != SAYAKA_FP64 && p->precision != SAYAKA_MIXED && p->precision != SAYAKA_FP64_FAST)
        return fail(-1, "bad precision");
    HIPCHK(hipSetDevice(c->device));
    return dispatch(c, [&](auto I) { return decltype(I)::solve(c, p, st); });
}

int sayaka_dev_copy(sayaka_dev_ctx* c, int dst, int src) {
    if (!c || !field_ok(dst) || !field_ok(src)) return fail(-1, "bad argument");
    HIPCHK(hipSetDevice(c->device));
    if (dst != src) {
        const size_t n2 = (size_t)c->nblk * c->ncell / 2;   // nb^d is even
        const size_t nwg = (n2 + 255) / 256;
        if (n2 > 0)
            hipLaunchKernelGGL(k_copy16, dim3((unsigned)nwg), dim3(256), 0, c->stream,
                               reinterpret_cast<const v2d*>(c->d_field[src]),
                               reinterpret_cast<v2d*>(c->d_field[dst]), n2);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int sayaka_dev_add(sayaka_dev_ctx* c, int m, int dst, int src) {
    if (!c || !field_ok(dst) || !field_ok(src) || m < 1 || m > c->finest) return fail(-1, "bad argument");
    HIPCHK(hipSetDevice(c->device));
    int e = dispatch_nc(c, [&](auto I) {
        using Im = decltype(I);
        VecArgs a = vec_args(c, c->tw[m]);
        a.dst = c->d_field[dst];
        a.x = c->d_field[src];
        return launch_vec<Im::D_, Im::NB_>(c, c->tw[m], V_FINAL, a, FIN_NONE);
    });
    if (e) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int sayaka_dev_fill(sayaka_dev_ctx* c, int field, double value) {
    if (!c || !field_ok(field)) return fail(-1, "bad argument");
    HIPCHK(hipSetDevice(c->device));
    if (value != 0.0) return fail(-1, "only zero fill is supported");
    int e = dispatch_nc(c, [&](auto I) {
        using Im = decltype(I);
        VecArgs a = vec_args(c, c->all_blocks);
        a.dst = c->d_field[field];
        return launch_vec<Im::D_, Im::NB_>(c, c->all_blocks, V_ZERO, a, FIN_NONE);
    });
    if (e) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int sayaka_dev_set_mixed_bottom(sayaka_dev_ctx* c, double tol_rel, double tol_abs, int max_iter) {
    if (!c || tol_rel < 0 || tol_abs < 0 || max_iter < 1) return fail(-1, "bad argument");
    c->mixed.bottom_tol_rel = tol_rel;
    c->mixed.bottom_tol_abs = tol_abs;
    c->mixed.bottom_max_iter = max_iter;
    return 0;
}

int sayaka_dev_set_mixed_coarse(sayaka_dev_ctx* c, int two_level) {
    if (!c) return fail(-1, "null ctx");
    c->pc_enable = two_level ? 1 : 0;
    c->pc_version = -1;
    return 0;
}

int sayaka_dev_set_mixed_smoother(sayaka_dev_ctx* c, int merged) {
    if (!c) return fail(-1, "null ctx");
    if (merged < 0 || merged > 2) return fail(-1, "smoother mode must be 0, 1 or 2");
    c->mixed.merged_smooth = merged;
    return 0;
}

int sayaka_dev_profile(sayaka_dev_ctx* c, int kclass, int enable) {
    if (!c || kclass < 0 || kclass >= SAYAKA_KCLASS_N) return fail(-1, "bad argument");
    HIPCHK(hipStreamSynchronize(c->stream));
    for (auto& r : c->prof[kclass]) {
        c->ev_pool.push_back(r.a);
        c->ev_pool.push_back(r.b);
    }
    c->prof[kclass].clear();
    c->prof_on[kclass] = enable != 0;
    return 0;
}

int sayaka_dev_profile_read(sayaka_dev_ctx* c, int kclass, double* total_ms, long long* launches, long long* cells,
                            double* alg_bytes) {
    if (!c || kclass < 0 || kclass >= SAYAKA_KCLASS_N) return fail(-1, "bad argument");
    HIPCHK(hipStreamSynchronize(c->stream));
    double t = 0.0, by = 0.0;
    long long n = 0, cl = 0;
    for (auto& r : c->prof[kclass]) {
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
        t += ms;
        n++;
        cl += r.cells;
        by += r.bytes;
    }
    if (total_ms) *total_ms = t;
    if (launches) *launches = n;
    if (cells) *cells = cl;
    if (alg_bytes) *alg_bytes = by;
    return 0;
}

int sayaka_dev_sync(sayaka_dev_ctx* c) {
    if (!c) return fail(-1, "null ctx");
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

}  // extern "C"

// ================================================================ learned initial guess
namespace {

constexpr int CNN_PL = cnn::NTAP * cnn::C * cnn::C + cnn::C;   // parameters per layer (padded)
constexpr int CNN_NP = cnn::NLAYER * CNN_PL;

void cnn_free(sayaka_dev_ctx::Cnn& n) {
    void* ptrs[] = {n.p, n.g, n.m, n.v, n.g0, n.g1, n.out32, n.tgt, n.loss, n.list, n.store, n.store_t};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    for (int l = 0; l < cnn::NLAYER; l++) {
        if (n.wf[l]) (void)hipFree(n.wf[l]);
        if (n.wb[l]) (void)hipFree(n.wb[l]);
        if (n.act[l]) (void)hipFree(n.act[l]);
    }
    n = sayaka_dev_ctx::Cnn{};
}

int cnn_pack(sayaka_dev_ctx* c) {
    for (int l = 0; l < cnn::NLAYER; l++)
        hipLaunchKernelGGL(cnn::k_cnn_pack, dim3(16), dim3(256), 0, c->stream, c->cnn.p + l * CNN_PL, c->cnn.wf[l],
                           c->cnn.wb[l]);
    HIPCHK(hipGetLastError());
    return 0;
}

int cnn_forward(sayaka_dev_ctx* c, int n) {
    auto& N = c->cnn;
    const dim3 g(n), b(cnn::NT);
    const float* bias[4];
    for (int l = 0; l < 4; l++) bias[l] = N.p + l * CNN_PL + cnn::NTAP * cnn::C * cnn::C;
    hipLaunchKernelGGL((cnn::k_conv_fwd<true, false>), g, b, 0, c->stream, N.act[0], N.wf[0], bias[0], N.act[1], nullptr);
    hipLaunchKernelGGL((cnn::k_conv_fwd<true, false>), g, b, 0, c->stream, N.act[1], N.wf[1], bias[1], N.act[2], nullptr);
    hipLaunchKernelGGL((cnn::k_conv_fwd<true, false>), g, b, 0, c->stream, N.act[2], N.wf[2], bias[2], N.act[3], nullptr);
    hipLaunchKernelGGL((cnn::k_conv_fwd<false, true>), g, b, 0, c->stream, N.act[3], N.wf[3], bias[3], nullptr, N.out32);
    HIPCHK(hipGetLastError());
    return 0;
}

// loss and parameter gradients of the last forward pass (targets in N.tgt)
int cnn_backward(sayaka_dev_ctx* c, int n) {
    auto& N = c->cnn;
    const dim3 g(n), b(cnn::NT);
    HIPCHK(hipMemsetAsync(N.g, 0, CNN_NP * sizeof(float), c->stream));
    HIPCHK(hipMemsetAsync(N.loss, 0, sizeof(float), c->stream));
    float* dw[4];
    float* db[4];
    for (int l = 0; l < 4; l++) {
        dw[l] = N.g + l * CNN_PL;
        db[l] = dw[l] + cnn::NTAP * cnn::C * cnn::C;
    }
    const float scale = 1.0f / ((float)n * 512.0f);
    hipLaunchKernelGGL(cnn::k_cnn_loss, g, b, 0, c->stream, N.out32, N.tgt, scale, N.g0, N.loss);
    // layer 4 (no activation): dP4 = dY4 (g0)
    hipLaunchKernelGGL((cnn::k_conv_bwd_w<false>), g, b, 0, c->stream, N.act[3], N.g0, N.g0, dw[3], db[3]);
    hipLaunchKernelGGL((cnn::k_conv_bwd_data<false>), g, b, 0, c->stream, N.g0, N.g0, N.wb[3], N.g1);
    // layer 3: dY3 in g1, Y3 = act[3]
    hipLaunchKernelGGL((cnn::k_conv_bwd_w<true>), g, b, 0, c->stream, N.act[2], N.g1, N.act[3], dw[2], db[2]);
    hipLaunchKernelGGL((cnn::k_conv_bwd_data<true>), g, b, 0, c->stream, N.g1, N.act[3], N.wb[2], N.g0);
    // layer 2: dY2 in g0, Y2 = act[2]
    hipLaunchKernelGGL((cnn::k_conv_bwd_w<true>), g, b, 0, c->stream, N.act[1], N.g0, N.act[2], dw[1], db[1]);
    hipLaunchKernelGGL((cnn::k_conv_bwd_data<true>), g, b, 0, c->stream, N.g0, N.act[2], N.wb[1], N.g1);
    // layer 1: dY1 in g1, Y1 = act[1]
    hipLaunchKernelGGL((cnn::k_conv_bwd_w<true>), g, b, 0, c->stream, N.act[0], N.g1, N.act[1], dw[0], db[0]);
    HIPCHK(hipGetLastError());
    return 0;
}

int cnn_fetch(sayaka_dev_ctx* c, const long long* idx, int n) {
    auto& N = c->cnn;
    for (int k = 0; k < n; k++)
        if (idx[k] < 0 || idx[k] >= N.store_n) return fail(-1, "training index out of range");
    long long* d_idx = reinterpret_cast<long long*>(N.list);   // list holds cap ints = cap/2 indices
    if (n > N.cap / 2) return fail(-1, "batch larger than half the window capacity");
    HIPCHK(hipMemcpyAsync(d_idx, idx, n * sizeof(long long), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(cnn::k_cnn_fetch, dim3(n), dim3(cnn::NT), 0, c->stream, N.store, N.store_t, d_idx, N.act[0],
                       N.tgt);
    HIPCHK(hipGetLastError());
    return 0;
}

}  // namespace

extern "C" {

int sayaka_cnn_setup(sayaka_dev_ctx* c, int max_windows) {
    if (!c || max_windows < 2) return fail(-1, "bad argument");
    if (c->ndim != 3 || c->nb != 8) return fail(-1, "the learned guess is defined for 8^3 blocks in 3D");
    HIPCHK(hipSetDevice(c->device));
    cnn_free(c->cnn);
    auto& N = c->cnn;
    N.cap = max_windows;
    const size_t na = (size_t)max_windows * cnn::NV * cnn::C;
    hipError_t e = hipSuccess;
    for (float** q : {&N.p, &N.g, &N.m, &N.v})
        if (e == hipSuccess) e = hipMalloc(q, CNN_NP * sizeof(float));
    for (int l = 0; l < cnn::NLAYER && e == hipSuccess; l++) {
        e = hipMalloc(&N.wf[l], cnn::WSZ * sizeof(__bf16));
        if (e == hipSuccess) e = hipMalloc(&N.wb[l], cnn::WSZ * sizeof(__bf16));
        if (e == hipSuccess) e = hipMalloc(&N.act[l], na * sizeof(__bf16));
    }
    if (e == hipSuccess) e = hipMalloc(&N.g0, na * sizeof(__bf16));
    if (e == hipSuccess) e = hipMalloc(&N.g1, na * sizeof(__bf16));
    if (e == hipSuccess) e = hipMalloc(&N.out32, (size_t)max_windows * cnn::NV * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&N.tgt, (size_t)max_windows * 512 * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&N.loss, sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&N.list, (size_t)max_windows * sizeof(int));
    for (float* q : {N.p, N.m, N.v})
        if (e == hipSuccess && q) e = hipMemset(q, 0, CNN_NP * sizeof(float));
    if (e != hipSuccess) {
        cnn_free(N);
        return fail(-3, std::string("cnn allocation: ") + hipGetErrorString(e));
    }
    return cnn_pack(c);
}

int sayaka_cnn_set_params(sayaka_dev_ctx* c, const float* params) {
    if (!c || !params || !c->cnn.p) return fail(-1, "cnn not set up");
    auto& N = c->cnn;
    HIPCHK(hipMemcpy(N.p, params, CNN_NP * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(N.m, 0, CNN_NP * sizeof(float)));
    HIPCHK(hipMemset(N.v, 0, CNN_NP * sizeof(float)));
    N.step = 0;
    int e = cnn_pack(c);
    if (e) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int sayaka_cnn_get_params(sayaka_dev_ctx* c, float* params) {
    if (!c || !params || !c->cnn.p) return fail(-1, "cnn not set up");
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(params, c->cnn.p, CNN_NP * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int sayaka_cnn_store_reset(sayaka_dev_ctx* c, long long cap) {
    if (!c || !c->cnn.p || cap < 1) return fail(-1, "bad argument");
    auto& N = c->cnn;
    if (N.store) (void)hipFree(N.store);
    if (N.store_t) (void)hipFree(N.store_t);
    N.store = nullptr;
    N.store_t = nullptr;
    N.store_n = N.store_cap = 0;
    HIPCHK(hipMalloc(&N.store, (size_t)cap * cnn::NV * cnn::C * sizeof(__bf16)));
    HIPCHK(hipMalloc(&N.store_t, (size_t)cap * 512 * sizeof(float)));
    N.store_cap = cap;
    return 0;
}

// windows of leaf blocks (host ids) of the current tree from RHS and the coefficients, targets
// (SOL - shift) * t_scale, appended to the training store; returns the store size
long long sayaka_cnn_collect(sayaka_dev_ctx* c, const int* leaf_ids, int n, double s_rhs, double shift,
                             double t_scale) {
    if (!c || !leaf_ids || n < 1 || !c->cnn.store) return fail(-1, "bad argument");
    if (c->coef_pending)
        if (int e = ensure_coefs(c)) return e;
    auto& N = c->cnn;
    if (N.store_n + n > N.store_cap) return fail(-1, "training store full");
    if (n > N.cap) return fail(-1, "more windows than the capacity");
    for (int k = 0; k < n; k++)
        if (leaf_ids[k] < 0 || leaf_ids[k] >= c->nblk || !c->h_binfo[leaf_ids[k]].leaf) return fail(-1, "not a leaf block");
    HIPCHK(hipMemcpyAsync(N.list, leaf_ids, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
    cnn::GatherArgs g{N.list, c->d_binfo, c->d_field[SAYAKA_RHS], c->d_b, c->d_af, c->d_vf, s_rhs};
    hipLaunchKernelGGL(cnn::k_cnn_gather, dim3(n), dim3(cnn::NT), 0, c->stream, g,
                       N.store + (size_t)N.store_n * cnn::NV * cnn::C);
    hipLaunchKernelGGL(cnn::k_cnn_target, dim3(n), dim3(256), 0, c->stream, N.list, c->d_field[SAYAKA_SOL], shift,
                       t_scale, N.store_t + (size_t)N.store_n * 512);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    N.store_n += n;
    return N.store_n;
}

int sayaka_cnn_train_step(sayaka_dev_ctx* c, const long long* idx, int n, float lr, float* loss) {
    if (!c || !idx || n < 1 || !c->cnn.p) return fail(-1, "bad argument");
    auto& N = c->cnn;
    int e;
    if ((e = cnn_fetch(c, idx, n)) || (e = cnn_forward(c, n)) || (e = cnn_backward(c, n))) return e;
    N.step++;
    cnn::AdamArgs a{N.p, N.g, N.m, N.v, CNN_NP, lr, 0.9f, 0.999f, 1e-8f,
                    1.0f - powf(0.9f, (float)N.step), 1.0f - powf(0.999f, (float)N.step)};
    hipLaunchKernelGGL(cnn::k_adam, dim3((CNN_NP + 255) / 256), dim3(256), 0, c->stream, a);
    HIPCHK(hipGetLastError());
    if ((e = cnn_pack(c))) return e;
    if (loss) {
        HIPCHK(hipMemcpyAsync(loss, N.loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        *loss /= (float)n * 512.0f;
    }
    return 0;
}

// loss and gradients (no update) of the store windows idx; grad has the parameter layout
int sayaka_cnn_loss_grad(sayaka_dev_ctx* c, const long long* idx, int n, float* loss, float* grad, float* out32) {
    if (!c || !idx || n < 1 || !c->cnn.p) return fail(-1, "bad argument");
    auto& N = c->cnn;
    int e;
    if ((e = cnn_fetch(c, idx, n)) || (e = cnn_forward(c, n)) || (e = cnn_backward(c, n))) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    float l = 0.0f;
    HIPCHK(hipMemcpy(&l, N.loss, sizeof(float), hipMemcpyDeviceToHost));
    if (loss) *loss = l / ((float)n * 512.0f);
    if (grad) HIPCHK(hipMemcpy(grad, N.g, CNN_NP * sizeof(float), hipMemcpyDeviceToHost));
    if (out32) HIPCHK(hipMemcpy(out32, N.out32, (size_t)n * cnn::NV * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

// copy store windows to the host (tests): [n][NV][C] as float
int sayaka_cnn_store_read(sayaka_dev_ctx* c, long long first, int n, float* windows, float* targets) {
    if (!c || !c->cnn.store || first < 0 || first + n > c->cnn.store_n) return fail(-1, "bad argument");
    auto& N = c->cnn;
    std::vector<__bf16> tmp((size_t)n * cnn::NV * cnn::C);
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(tmp.data(), N.store + (size_t)first * cnn::NV * cnn::C, tmp.size() * sizeof(__bf16),
                     hipMemcpyDeviceToHost));
    if (windows)
        for (size_t q = 0; q < tmp.size(); q++) windows[q] = (float)tmp[q];
    if (targets)
        HIPCHK(hipMemcpy(targets, N.store_t + (size_t)first * 512, (size_t)n * 512 * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

// initial guess of every leaf block of the finest tower into `field` (other blocks zero)
int sayaka_cnn_guess(sayaka_dev_ctx* c, int field, double s_rhs, double out_scale) {
    if (!c || !c->cnn.p || !field_ok(field)) return fail(-1, "bad argument");
    auto& N = c->cnn;
    std::vector<int> leaves;
    for (int b = 0; b < c->nblk; b++)
        if (c->h_binfo[b].leaf) leaves.push_back(b);
    int e = dispatch(c, [&](auto I) {
        using Im = decltype(I);
        VecArgs a = vec_args(c, c->all_blocks);
        a.dst = c->d_field[field];
        return launch_vec<Im::D_, Im::NB_>(c, c->all_blocks, V_ZERO, a, FIN_NONE);
    });
    if (e) return e;
    for (size_t off = 0; off < leaves.size(); off += N.cap) {
        const int n = (int)std::min<size_t>(N.cap, leaves.size() - off);
        HIPCHK(hipMemcpyAsync(N.list, leaves.data() + off, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
        cnn::GatherArgs g{N.list, c->d_binfo, c->d_field[SAYAKA_RHS], c->d_b, c->d_af, c->d_vf, s_rhs};
        hipLaunchKernelGGL(cnn::k_cnn_gather, dim3(n), dim3(cnn::NT), 0, c->stream, g, N.act[0]);
        if ((e = cnn_forward(c, n))) return e;
        hipLaunchKernelGGL(cnn::k_cnn_emit, dim3(n), dim3(256), 0, c->stream, N.list, N.out32, c->d_vf, out_scale,
                           c->d_field[field]);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));   // N.list is reused by the next batch
    }
    return 0;
}

// generator scratch for the current tree: block corners [nblk][3], sizes [nblk], partial sums
// [2 nblk]; lo3 (and dx when given) uploaded
static int gen_scratch(sayaka_dev_ctx* c, const double* lo3, const double* dx) {
    const int n = c->nblk;
    if (c->gen_cap < n) {
        if (c->d_gen) (void)hipFree(c->d_gen);
        c->d_gen = nullptr;
        HIPCHK(hipMalloc(&c->d_gen, (size_t)n * 6 * sizeof(double)));
        c->gen_cap = n;
    }
    HIPCHK(hipMemcpyAsync(c->d_gen, lo3, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (dx) HIPCHK(hipMemcpyAsync(c->d_gen + 3 * n, dx, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    return 0;
}

// cut-cell fractions and B of the solid sphere, coarse/fine faces synchronised, weights
static int gen_sphere_coefs(sayaka_dev_ctx* c, const double* lo3, const double* dx, const double* centre,
                            double radius, double ratio, double eps, int nsub) {
    if (c->ndim != 3) return fail(-1, "3D only");
    if (c->nb != 8 && c->nb != 16) return fail(-1, "8^3 or 16^3 blocks");
    HIPCHK(hipSetDevice(c->device));
    int e = gen_scratch(c, lo3, dx);
    if (e) return e;
    const int n = c->nblk;
    if (nsub < 1 || nsub > cnn::MAXSUB) return fail(-1, "nsub must be 1 .. 8");
    cnn::SphereArgs sa{c->d_gen, c->d_gen + 3 * n, {centre[0], centre[1], centre[2]}, radius, ratio, eps, nsub};
    c->coef_pending = 0;   // the generator writes vf, a_f and B of every block
    if (c->nb == 8)
        hipLaunchKernelGGL(cnn::k_gen_sphere<8>, dim3(n), dim3(256), 0, c->stream, sa, c->d_vf, c->d_af, c->d_b);
    else
        hipLaunchKernelGGL(cnn::k_gen_sphere<16>, dim3(n), dim3(256), 0, c->stream, sa, c->d_vf, c->d_af, c->d_b);
    HIPCHK(hipGetLastError());
    e = dispatch(c, [&](auto I) { return decltype(I)::set_coefs(c); });
    if (e) return e;
    c->coef_version++;
    return 0;
}

// Inputs of a solid sphere in a density-1 fluid (cut-cell fractions, B = 1/rho) and a
// Gaussian-blob rhs projected to zero mean over the fluid leaves, generated on the device for
// the current tree (lo3 [nblk][3], dx [nblk] from the tree).  Returns the removed mean.
int sayaka_gen_sphere_blobs(sayaka_dev_ctx* c, const double* lo3, const double* dx, const double* centre, double radius,
                            double ratio, double eps, int nsub, int nblob, const double* blob_c, const double* blob_w,
                            const double* blob_a, double* mean_out) {
    if (!c || !lo3 || !dx || !centre || nsub < 1 || nblob < 0 || nblob > cnn::MAXBLOB) return fail(-1, "bad argument");
    int e = gen_sphere_coefs(c, lo3, dx, centre, radius, ratio, eps, nsub);
    if (e) return e;
    const int n = c->nblk;
    double* d_part = c->d_gen + 4 * n;
    cnn::BlobArgs ba{};
    ba.lo = c->d_gen;
    ba.dx = c->d_gen + 3 * n;
    ba.nblob = nblob;
    for (int q = 0; q < nblob; q++) {
        for (int d = 0; d < 3; d++) ba.c[q][d] = blob_c[3 * q + d];
        ba.w[q] = blob_w[q];
        ba.amp[q] = blob_a[q];
    }
    double* rhs = c->d_field[SAYAKA_RHS];
    if (c->nb == 8)
        hipLaunchKernelGGL(cnn::k_gen_blobs<8>, dim3(n), dim3(256), 0, c->stream, ba, c->d_binfo, c->d_vf, rhs, d_part);
    else
        hipLaunchKernelGGL(cnn::k_gen_blobs<16>, dim3(n), dim3(256), 0, c->stream, ba, c->d_binfo, c->d_vf, rhs, d_part);
    HIPCHK(hipGetLastError());
    std::vector<double> part(2 * n);
    HIPCHK(hipMemcpyAsync(part.data(), d_part, 2 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double s = 0.0, w = 0.0;
    for (int b = 0; b < n; b++) {
        s += part[2 * b];
        w += part[2 * b + 1];
    }
    const double mean = w > 0 ? s / w : 0.0;
    const int ncell = n * c->ncell;
    hipLaunchKernelGGL(cnn::k_sub_mean, dim3(1024), dim3(256), 0, c->stream, rhs, c->d_vf, mean, ncell);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    if (mean_out) *mean_out = mean;
    return 0;
}

int sayaka_gen_sphere(sayaka_dev_ctx* c, const double* lo3, const double* dx, const double* centre, double radius,
                      double ratio, double eps, int nsub) {
    if (!c || !lo3 || !dx || !centre || nsub < 1) return fail(-1, "bad argument");
    int e = gen_sphere_coefs(c, lo3, dx, centre, radius, ratio, eps, nsub);
    if (e) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

static int vel_alloc(sayaka_dev_ctx* c) {
    if (!c->d_vel) HIPCHK(hipMalloc(&c->d_vel, (size_t)c->cap * c->ndim * c->nface * sizeof(double)));
    if (!c->d_div) HIPCHK(hipMalloc(&c->d_div, (size_t)c->cap * c->ncell * sizeof(double)));
    return 0;
}

int sayaka_gen_velocity(sayaka_dev_ctx* c, const double* lo3, const double* dlo3, const double* root3,
                        unsigned long long seed, int zero_walls) {
    if (!c || !lo3 || !dlo3 || !root3 || c->nblk < 1) return fail(-1, "bad argument");
    for (int d = 0; d < c->ndim; d++)
        if (!(root3[d] > 0)) return fail(-1, "root block edge must be positive");
    HIPCHK(hipSetDevice(c->device));
    int e = vel_alloc(c);
    if (!e) e = gen_scratch(c, lo3, nullptr);
    if (!e)
        e = dispatch_nc(c, [&](auto I) {
            return decltype(I)::gen_velocity(c, c->d_gen, dlo3, root3, seed, zero_walls ? 1 : 0);
        });
    if (e) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    c->vel_valid = 1;
    return 0;
}

static int need_vel(sayaka_dev_ctx* c) {
    if (!c->vel_valid) return fail(-1, "no face velocity for this tree: call sayaka_gen_velocity");
    return 0;
}

int sayaka_gen_divergence_rhs(sayaka_dev_ctx* c, double* mean_out) {
    if (!c) return fail(-1, "bad argument");
    int e = need_vel(c);
    if (e) return e;
    HIPCHK(hipSetDevice(c->device));
    e = dispatch(c, [&](auto I) {
        using Im = decltype(I);
        Im::sync_faces(c, c->d_vel);
        return Im::divergence(c, c->d_field[SAYAKA_RHS], c->d_gen + 4 * c->nblk, mean_out);
    });
    if (e) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int sayaka_dev_project(sayaka_dev_ctx* c, int phi, double* div_before, double* div_after) {
    if (!c || !field_ok(phi) || c->finest < 1) return fail(-1, "bad argument");
    int e = need_vel(c);
    if (e) return e;
    HIPCHK(hipSetDevice(c->device));
    double b = 0, a = 0;
    e = dispatch(c, [&](auto I) {
        return decltype(I)::project(c, c->d_field[phi], c->d_gen + 4 * c->nblk, &b, &a);
    });
    if (e) return e;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (div_before) *div_before = b;
    if (div_after) *div_after = a;
    return 0;
}

int sayaka_dev_get_velocity(sayaka_dev_ctx* c, double* out) {
    if (!c || !out) return fail(-1, "bad argument");
    int e = need_vel(c);
    if (e) return e;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(out, c->d_vel, (size_t)c->nblk * c->ndim * c->nface * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

}  // extern "C"

#ifdef SAYAKA_BOTTOM_PHASES
// Phase times of the coarsest-level solve since the last reset (the phase-timing build only):
// out32[k] wall-clock ticks of phase k (sydev::BottomPhase), [30] iterations, [31] launches.
extern "C" int sayaka_dev_bottom_phases(int device, unsigned long long* out32, int reset, int* clock_khz) {
    if (!out32) return fail(-1, "bad argument");
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpyFromSymbol(out32, HIP_SYMBOL(sydev::g_bottom_phases), 32 * sizeof(unsigned long long)));
    if (clock_khz) HIPCHK(hipDeviceGetAttribute(clock_khz, hipDeviceAttributeWallClockRate, device));
    if (reset) {
        const unsigned long long z[32] = {};
        HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(sydev::g_bottom_phases), z, sizeof(z)));
    }
    return 0;
}
#endif
