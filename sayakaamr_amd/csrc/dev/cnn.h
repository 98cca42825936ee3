// cnn.h -- the learned initial guess: a per-block 3D CNN trained and run on the GPU.
//
// Input: one 12^3 window per leaf block (its 8^3 cells plus two cells of same-level
// neighbours on every face; edges, corners and non-same-level neighbours are zero), CIN = 16
// channels of which 5 are used (scaled rhs, three cell-averaged B*a face coefficients, the
// volume fraction).  Net: four 3x3x3 'same' convolutions 16 -> 16 -> 16 -> 16 -> 1 (the first
// reads the 5 used input channels, zero-padded to 16; the last writes one channel), ReLU after
// the first three.  Output: the central 8^3 of channel 0, the initial guess of the block.
//
// Arithmetic: activations and weights in BF16, products summed in FP32 on the matrix cores
// (v_mfma_f32_16x16x32_bf16), the last layer's output kept in FP32.  Every convolution is an
// implicit GEMM over one block's window staged in LDS:
//   forward      Y[v][co]   = sum_{tap, ci} X[v + off(tap)][ci] * W[tap][ci][co]   (M = voxels)
//   data grad    dX[v][ci]  = sum_{tap, co} dP[v - off(tap)][co] * W[tap][ci][co]  (M = voxels)
//   weight grad  dW[tap][ci][co] = sum_v X[v + off(tap)][ci] * dP[v][co]           (K = voxels)
// with dP = dY masked by the ReLU (Y > 0).  K = 27 taps x 16 channels runs in 14 steps of 32
// (two taps per step, the 28th tap zero).
//
// MFMA lane layout (16x16x32 BF16, checked on gfx950): lane l holds A[l % 16][(l / 16) * 8 + q]
// and B[(l / 16) * 8 + q][l % 16] for q < 8, and receives C[(l / 16) * 4 + q][l % 16], q < 4.
#pragma once

#include <hip/hip_runtime.h>

namespace sydev {

namespace cnn {

constexpr int WIN = 12, NV = WIN * WIN * WIN;   // window edge, voxels
constexpr int C = 16;                            // channels of every activation buffer
constexpr int NTAP = 27, NTAPP = 28;             // taps, padded to an even count
constexpr int NLAYER = 4;
constexpr int NT = 256;                          // threads per workgroup (4 waves)
constexpr int NTILE = NV / 16;                   // 108 voxel tiles of 16
constexpr int WSZ = NTAPP * C * C;               // one layer's BF16 weights (padded)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline void tap_off(int tap, int& dx, int& dy, int& dz) {
    dx = tap % 3 - 1;
    dy = (tap / 3) % 3 - 1;
    dz = tap / 9 - 1;
}
__device__ inline int vox(int x, int y, int z) { return x + WIN * (y + WIN * z); }
__device__ inline bool inwin(int x, int y, int z) {
    return (unsigned)x < (unsigned)WIN && (unsigned)y < (unsigned)WIN && (unsigned)z < (unsigned)WIN;
}

// 16-byte copies global -> LDS of n bf16 values (n a multiple of 8)
__device__ inline void copy_bf16(__bf16* dst, const __bf16* src, int n) {
    const uint4* s = reinterpret_cast<const uint4*>(src);
    uint4* d = reinterpret_cast<uint4*>(dst);
    for (int q = threadIdx.x; q < n / 8; q += blockDim.x) d[q] = s[q];
}

// ------------------------------------------------------------------------------- forward
// in:  [blk][NV][C] BF16; wf: [NTAPP][co][ci] BF16 (tap 27 zero); bias [C] FP32.
// out: [blk][NV][C] BF16 (post-activation) or, for the last layer, out32 [blk][NV] FP32 (co 0).
template <bool RELU, bool LAST>
__global__ void __launch_bounds__(NT) k_conv_fwd(const __bf16* __restrict__ in, const __bf16* __restrict__ wf,
                                                  const float* __restrict__ bias, __bf16* __restrict__ out,
                                                  float* __restrict__ out32) {
    __shared__ __attribute__((aligned(16))) __bf16 xs[NV * C];
    __shared__ __attribute__((aligned(16))) __bf16 ws[WSZ];
    const int blk = blockIdx.x;
    copy_bf16(xs, in + (size_t)blk * NV * C, NV * C);
    copy_bf16(ws, wf, WSZ);
    __syncthreads();
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int half = l >> 5, cb = ((l >> 4) & 1) * 8, col = l & 15;
    const float bv = bias[col];
    for (int tile = wave; tile < NTILE; tile += NT / 64) {
        const int v = tile * 16 + (l & 15);
        const int x = v % WIN, y = (v / WIN) % WIN, z = v / (WIN * WIN);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kc = 0; kc < NTAPP / 2; kc++) {
            const int tap = 2 * kc + half;
            int dx, dy, dz;
            tap_off(tap, dx, dy, dz);
            bf16x8 a;
            if (tap < NTAP && inwin(x + dx, y + dy, z + dz))
                a = *reinterpret_cast<const bf16x8*>(xs + vox(x + dx, y + dy, z + dz) * C + cb);
            else
                for (int q = 0; q < 8; q++) a[q] = (__bf16)0.0f;
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(ws + (tap * C + col) * C + cb);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
        for (int q = 0; q < 4; q++) {
            const int vo = tile * 16 + (l >> 4) * 4 + q;
            float r = acc[q] + bv;
            if (RELU) r = fmaxf(r, 0.0f);
            if (LAST) {
                if (col == 0) out32[(size_t)blk * NV + vo] = r;
            } else {
                out[((size_t)blk * NV + vo) * C + col] = (__bf16)r;
            }
        }
    }
}

// dP = dY masked by the ReLU of Y (both [blk][NV][C]); staged into LDS [NV][C]
__device__ inline void stage_dp(__bf16* dst, const __bf16* dy, const __bf16* y, bool relu) {
    for (int q = threadIdx.x; q < NV * C; q += blockDim.x) {
        __bf16 g = dy[q];
        if (relu && !((float)y[q] > 0.0f)) g = (__bf16)0.0f;
        dst[q] = g;
    }
}

// ---------------------------------------------------------------------------- data grad
// dX[v][ci] = sum_{tap, co} dP[v - off(tap)][co] * W[tap][ci][co]; wb: [NTAPP][ci][co] BF16.
template <bool RELU>
__global__ void __launch_bounds__(NT) k_conv_bwd_data(const __bf16* __restrict__ dy, const __bf16* __restrict__ y,
                                                       const __bf16* __restrict__ wb, __bf16* __restrict__ dx) {
    __shared__ __attribute__((aligned(16))) __bf16 ps[NV * C];
    __shared__ __attribute__((aligned(16))) __bf16 ws[WSZ];
    const int blk = blockIdx.x;
    stage_dp(ps, dy + (size_t)blk * NV * C, y + (size_t)blk * NV * C, RELU);
    copy_bf16(ws, wb, WSZ);
    __syncthreads();
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int half = l >> 5, cb = ((l >> 4) & 1) * 8, col = l & 15;
    for (int tile = wave; tile < NTILE; tile += NT / 64) {
        const int v = tile * 16 + (l & 15);
        const int x = v % WIN, yy = (v / WIN) % WIN, z = v / (WIN * WIN);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kc = 0; kc < NTAPP / 2; kc++) {
            const int tap = 2 * kc + half;
            int ox, oy, oz;
            tap_off(tap, ox, oy, oz);
            bf16x8 a;
            if (tap < NTAP && inwin(x - ox, yy - oy, z - oz))
                a = *reinterpret_cast<const bf16x8*>(ps + vox(x - ox, yy - oy, z - oz) * C + cb);
            else
                for (int q = 0; q < 8; q++) a[q] = (__bf16)0.0f;
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(ws + (tap * C + col) * C + cb);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
        for (int q = 0; q < 4; q++) {
            const int vo = tile * 16 + (l >> 4) * 4 + q;
            dx[((size_t)blk * NV + vo) * C + col] = (__bf16)acc[q];
        }
    }
}

// -------------------------------------------------------------------------- weight grad
// dW[tap][ci][co] += sum_v X[v + off(tap)][ci] * dP[v][co] and db[co] += sum_v dP[v][co],
// summed over the workgroup's block in FP32 and added to the global FP32 gradients.
// LDS holds X and dP transposed ([channel][voxel]) so that 8 consecutive voxels are one load.
template <bool RELU>
__global__ void __launch_bounds__(NT) k_conv_bwd_w(const __bf16* __restrict__ xin, const __bf16* __restrict__ dy,
                                                    const __bf16* __restrict__ y, float* __restrict__ dw,
                                                    float* __restrict__ db) {
    __shared__ __attribute__((aligned(16))) __bf16 xt[C * NV];
    __shared__ __attribute__((aligned(16))) __bf16 pt[C * NV];
    const int blk = blockIdx.x;
    const __bf16* xb = xin + (size_t)blk * NV * C;
    const __bf16* dyb = dy + (size_t)blk * NV * C;
    const __bf16* yb = y + (size_t)blk * NV * C;
    for (int q = threadIdx.x; q < NV * C; q += NT) {
        const int v = q / C, c = q % C;
        xt[c * NV + v] = xb[q];
        __bf16 g = dyb[q];
        if (RELU && !((float)yb[q] > 0.0f)) g = (__bf16)0.0f;
        pt[c * NV + v] = g;
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int row = l & 15, kq = (l >> 4) * 8, col = l & 15;
    for (int tap = wave; tap < NTAP; tap += NT / 64) {
        int ox, oy, oz;
        tap_off(tap, ox, oy, oz);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < NV; k0 += 32) {
            bf16x8 a, b;
            for (int q = 0; q < 8; q++) {
                const int v = k0 + kq + q;
                const int x = v % WIN, yy = (v / WIN) % WIN, z = v / (WIN * WIN);
                a[q] = inwin(x + ox, yy + oy, z + oz) ? xt[row * NV + vox(x + ox, yy + oy, z + oz)] : (__bf16)0.0f;
            }
            b = *reinterpret_cast<const bf16x8*>(pt + col * NV + k0 + kq);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
        // C[m = ci][n = co]: lane holds rows (l / 16) * 4 + q, column l % 16
        for (int q = 0; q < 4; q++) {
            const int ci = (l >> 4) * 4 + q;
            atomicAdd(&dw[(tap * C + ci) * C + col], acc[q]);
        }
    }
    // bias gradient: sum over voxels of dP[.][co] (16 partial sums per channel, then LDS)
    __shared__ float bred[NT];
    {
        const int c = threadIdx.x & (C - 1), part = threadIdx.x / C;
        float sacc = 0.0f;
        for (int v = part; v < NV; v += NT / C) sacc += (float)pt[c * NV + v];
        bred[threadIdx.x] = sacc;
    }
    __syncthreads();
    if (threadIdx.x < C) {
        float sacc = 0.0f;
        for (int part = 0; part < NT / C; part++) sacc += bred[part * C + threadIdx.x];
        atomicAdd(&db[threadIdx.x], sacc);
    }
}

// ------------------------------------------------------------------------- loss and Adam
// MSE over the central 8^3 of channel 0 against the target [blk][512]; writes dY (BF16,
// [blk][NV][C], zero outside the centre and for channels > 0) and adds the loss sum.
__global__ void __launch_bounds__(NT) k_cnn_loss(const float* __restrict__ out32, const float* __restrict__ target,
                                                  float scale, __bf16* __restrict__ dy, float* __restrict__ loss) {
    const int blk = blockIdx.x;
    __shared__ float red[NT];
    float s = 0.0f;
    for (int q = threadIdx.x; q < NV * C; q += NT) dy[(size_t)blk * NV * C + q] = (__bf16)0.0f;
    __syncthreads();
    for (int c = threadIdx.x; c < 512; c += NT) {
        const int i = c % 8, j = (c / 8) % 8, k = c / 64;
        const int v = vox(i + 2, j + 2, k + 2);
        const float e = out32[(size_t)blk * NV + v] - target[(size_t)blk * 512 + c];
        s += e * e;
        dy[((size_t)blk * NV + v) * C] = (__bf16)(2.0f * e * scale);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = NT / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// Adam on the FP32 master parameters (one thread per parameter).  Parameter layout per layer:
// weights [27][ci][co] (padded to 16 x 16; the padded entries never receive a gradient, so
// they stay zero) then bias [16], for the four layers in order.
struct AdamArgs {
    float* p;
    const float* g;
    float* m;
    float* v;
    int n;
    float lr, b1, b2, eps, bc1, bc2;   // bc = 1 - beta^t
};
__global__ void k_adam(AdamArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const float g = a.g[i];
    const float m = a.b1 * a.m[i] + (1.0f - a.b1) * g;
    const float v = a.b2 * a.v[i] + (1.0f - a.b2) * g * g;
    a.m[i] = m;
    a.v[i] = v;
    const float mh = m / a.bc1, vh = v / a.bc2;
    a.p[i] -= a.lr * mh / (sqrtf(vh) + a.eps);
}

// FP32 master weights of one layer, [27][ci][co] padded to 16 x 16 -> the BF16 forward
// ([tap][co][ci]) and data-grad ([tap][ci][co]) copies, tap 27 zero.
__global__ void k_cnn_pack(const float* __restrict__ w, __bf16* __restrict__ wf, __bf16* __restrict__ wb) {
    for (int q = threadIdx.x + blockIdx.x * blockDim.x; q < WSZ; q += blockDim.x * gridDim.x) {
        const int tap = q / (C * C), a = (q / C) % C, b = q % C;
        wf[q] = tap < NTAP ? (__bf16)w[(tap * C + b) * C + a] : (__bf16)0.0f;
        wb[q] = tap < NTAP ? (__bf16)w[(tap * C + a) * C + b] : (__bf16)0.0f;
    }
}

}  // namespace cnn
}  // namespace sydev
