// Radiance cache: an MLP 81 -> 64 -> 64 -> 64 -> 64 -> 3 (ReLU hidden, exp
// output, 17,923 parameters) evaluated and trained with BF16 MFMA
// (v_mfma_f32_16x16x32_bf16, two 16-wide k-tiles per product, and
// v_mfma_f32_16x16x16_bf16 for single tiles) and FP32 accumulation.  No
// counterpart in the reference.
//
// Every GEMM is computed transposed, D = W . X^T, with the 16 samples of a
// wave as the N dimension: the D layout of a 16x16 output tile (lane l holds
// column l % 16 = sample and rows 4 (l / 16) + e = features) is exactly the
// B-operand layout of the next layer's k-tile, so activations move from layer
// to layer in registers without shuffles.  Weights live in LDS as packed A
// tiles (lane l: row l % 16, k = 4 (l / 16) + e).
//
// Training runs in four kernels: (1) forward + backward per wave of 16
// samples, writing the BF16 layer inputs and pre-activation gradients in a
// sample-quad interleaved layout; (2) the weight gradients dW = sum dz (x) a as
// split-K MFMA tiles over the samples; (3) a fixed-order reduction of the
// split partials plus the bias gradients; (4) Adam on the FP32 master weights
// and re-packing of the BF16 tiles.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "rc_internal.hpp"

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kIn = 81, kInPad = 96, kH = 64, kOut = 3, kLayers = 5;
constexpr int kParams = 17923;
static_assert(kParams == RC_CACHE_PARAMS, "parameter count");
// FP32 master layout (the oracle's): per layer W (out x in, row-major), b
constexpr int kOffW[kLayers] = {0, 5248, 9408, 13568, 17728};
constexpr int kOffB[kLayers] = {5184, 9344, 13504, 17664, 17920};
__host__ __device__ constexpr int fan_in(int l) { return l == 0 ? kIn : kH; }
__host__ __device__ constexpr int fan_out(int l) { return l == 4 ? kOut : kH; }
__host__ __device__ constexpr int ktiles(int l) { return l == 0 ? 6 : 4; }
__host__ __device__ constexpr int ntiles(int l) { return l == 4 ? 1 : 4; }
// forward A tiles: layer l, out tile nt, in tile kt
constexpr int kFwdBase[kLayers] = {0, 24, 40, 56, 72};
constexpr int kFwdTiles = 76;
__host__ __device__ constexpr int fwd_tile(int l, int nt, int kt) { return kFwdBase[l] + nt * ktiles(l) + kt; }
// backward A tiles (W^T) for layers 1..4: in tile nt (4), out tile kt
constexpr int kBwdBase[kLayers] = {0, 0, 16, 32, 48};
constexpr int kBwdTiles = 52;
__host__ __device__ constexpr int bwd_tile(int l, int nt, int kt) { return kBwdBase[l] + nt * ntiles(l) + kt; }
// per-sample BF16 storage (features): layer inputs a_0..a_4 and gradients dz_0..dz_4
constexpr int kActF = kInPad + 4 * kH;  // 352: x(96) h1 h2 h3 h4
constexpr int kActBase[kLayers] = {0, 96, 160, 224, 288};
constexpr int kDzF = 4 * kH + 16;  // 272: dz0..dz3 (64 each), dz4 (16)
constexpr int kDzBase[kLayers] = {0, 64, 128, 192, 256};
constexpr int kWgTiles = 76;  // weight-gradient tiles: (l, ot, it) in forward-tile order
constexpr int kSplits = 64;
constexpr int kThreads = 256;

struct Bounds {
    float lo[3];
    float inv[3];
};

__device__ __forceinline__ short f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return short(u >> 16);
}
__device__ __forceinline__ float bf2f(short h) { return __uint_as_float(uint32_t(uint16_t(h)) << 16); }
__device__ __forceinline__ v4s pack4(v4f x) {
    v4s r;
    r.x = f2bf(x.x);
    r.y = f2bf(x.y);
    r.z = f2bf(x.z);
    r.w = f2bf(x.w);
    return r;
}
__device__ __forceinline__ v4f relu4(v4f x) {
    return v4f{fmaxf(x.x, 0.0f), fmaxf(x.y, 0.0f), fmaxf(x.z, 0.0f), fmaxf(x.w, 0.0f)};
}
__device__ __forceinline__ v4f mfma(v4s a, v4s b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// Two k-tiles in one v_mfma_f32_16x16x32_bf16 (gfx950; twice the work of the
// 16x16x16 form per issue): each lane passes its values of k-tile 0 followed
// by those of k-tile 1 in both operands.  A and B put the same (lane group,
// element) at the same k index, so the product sums every k of both tiles
// once, whatever k each (lane group, element) stands for.
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef short v8s __attribute__((ext_vector_type(8)));
__device__ __forceinline__ v4f mfma2(v4s a0, v4s a1, v4s b0, v4s b1, v4f c) {
    const v8s a = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
    const v8s b = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c, 0, 0,
                                                   0);
}

// Input k-tile kt of the 96-wide padded encoding for this lane: features
// f = 16 kt + 4 (lane / 16) + e.  Features 0..71 are sin/cos(2^k pi p_a) at
// f = 24 a + 2 k (+1 for cos); a group of four starts at a multiple of 4, so it
// is (sin, cos) of two consecutive frequencies of one axis.  Then direction,
// normal, albedo (72..80) and zeros.
__device__ __forceinline__ v4s encode_tile(int kt, const float p01[3], const float* q, bool valid, int lane) {
    const int f0 = 16 * kt + 4 * (lane >> 4);
    v4f v = {0.0f, 0.0f, 0.0f, 0.0f};
    if (valid) {
        if (f0 < 72) {
            int a = f0 / 24, k = (f0 % 24) >> 1;
            float s0, c0, s1, c1;
            sincospif(ldexpf(p01[a], k), &s0, &c0);
            sincospif(ldexpf(p01[a], k + 1), &s1, &c1);
            v = v4f{s0, c0, s1, c1};
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) v[e] = f0 + e < kIn ? q[3 + (f0 + e - 72)] : 0.0f;
        }
    }
    return pack4(v);
}

__device__ __forceinline__ v4f bias_tile(const float* __restrict__ params, int l, int nt, int lane) {
    v4f b;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        int o = 16 * nt + 4 * (lane >> 4) + e;
        b[e] = o < fan_out(l) ? params[kOffB[l] + o] : 0.0f;
    }
    return b;
}

// Layer 0: the encoding is produced one input k-tile at a time and consumed
// by the four output tiles at once, so only one input tile is live.  When
// x_out is given, the input tiles are also written there (training).
__device__ __forceinline__ void layer0(const v4s* __restrict__ wt, const float* __restrict__ params, int lane,
                                       const float p01[3], const float* q, bool valid, v4s h0[4], v4s* x_out) {
    v4f acc[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) acc[nt] = bias_tile(params, 0, nt, lane);
#pragma unroll
    for (int kt = 0; kt < 6; kt += 2) {
        __builtin_amdgcn_sched_barrier(0);  // keep one k-step of weights in flight
        const v4s x0 = encode_tile(kt, p01, q, valid, lane), x1 = encode_tile(kt + 1, p01, q, valid, lane);
        if (x_out) {
            x_out[kt] = x0;
            x_out[kt + 1] = x1;
        }
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
            acc[nt] = mfma2(wt[fwd_tile(0, nt, kt) * 64 + lane], wt[fwd_tile(0, nt, kt + 1) * 64 + lane], x0, x1,
                            acc[nt]);
    }
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) h0[nt] = pack4(relu4(acc[nt]));
}

__device__ __forceinline__ void hidden(const v4s* __restrict__ wt, const float* __restrict__ params, int lane, int l,
                                       const v4s in[4], v4s out[4]) {
    v4f acc[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) acc[nt] = bias_tile(params, l, nt, lane);
#pragma unroll
    for (int kt = 0; kt < 4; kt += 2) {
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
            acc[nt] = mfma2(wt[fwd_tile(l, nt, kt) * 64 + lane], wt[fwd_tile(l, nt, kt + 1) * 64 + lane], in[kt],
                            in[kt + 1], acc[nt]);
    }
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) out[nt] = pack4(relu4(acc[nt]));
}

__device__ __forceinline__ v4f output_layer(const v4s* __restrict__ wt, const float* __restrict__ params, int lane,
                                            const v4s in[4]) {
    v4f acc = bias_tile(params, 4, 0, lane);
#pragma unroll
    for (int kt = 0; kt < 4; kt += 2) {
        __builtin_amdgcn_sched_barrier(0);
        acc = mfma2(wt[fwd_tile(4, 0, kt) * 64 + lane], wt[fwd_tile(4, 0, kt + 1) * 64 + lane], in[kt], in[kt + 1],
                    acc);
    }
    return acc;
}

__device__ __forceinline__ void load_tiles(v4s* dst, const v4s* __restrict__ src, int n) {
    for (int i = threadIdx.x; i < n; i += kThreads) dst[i] = src[i];
}

// Query record formats: API queries are 12 floats (pos, dir, normal, albedo);
// renderer queries are 4 float4: (pos, slot), (dir, thr.r), (normal, thr.g),
// (albedo, thr.b).
template <bool RENDER>
__global__ __launch_bounds__(kThreads) void k_cache_infer(const v4s* __restrict__ wt_g,
                                                          const float* __restrict__ params, Bounds B,
                                                          const float* __restrict__ queries, uint32_t n_host,
                                                          const uint32_t* __restrict__ n_dev, float* __restrict__ out) {
    __shared__ v4s wt[kFwdTiles * 64];
    load_tiles(wt, wt_g, kFwdTiles * 64);
    __syncthreads();
    const uint32_t n = n_dev ? min(*n_dev, n_host) : n_host;
    const int lane = threadIdx.x & 63;
    const uint32_t wave = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    const uint32_t nwaves = gridDim.x * (kThreads / 64);
    for (uint32_t g = wave; g * 16 < n; g += nwaves) {
        uint32_t qi = g * 16 + (lane & 15);
        bool valid = qi < n;
        float q[12];
        float thr[3] = {0.0f, 0.0f, 0.0f};
        uint32_t slot = 0;
        if (RENDER) {
            const float4* r = reinterpret_cast<const float4*>(queries) + 4 * size_t(valid ? qi : 0);
            float4 a = r[0], b = r[1], c = r[2], d = r[3];
            q[0] = a.x; q[1] = a.y; q[2] = a.z; slot = __float_as_uint(a.w);
            q[3] = b.x; q[4] = b.y; q[5] = b.z; thr[0] = b.w;
            q[6] = c.x; q[7] = c.y; q[8] = c.z; thr[1] = c.w;
            q[9] = d.x; q[10] = d.y; q[11] = d.z; thr[2] = d.w;
        } else {
#pragma unroll
            for (int k = 0; k < 12; ++k) q[k] = valid ? queries[size_t(qi) * 12 + k] : 0.0f;
        }
        float p01[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) p01[a] = (q[a] - B.lo[a]) * B.inv[a];
        v4s ha[4], hb[4];
        layer0(wt, params, lane, p01, q, valid, ha, nullptr);
        hidden(wt, params, lane, 1, ha, hb);
        hidden(wt, params, lane, 2, hb, ha);
        hidden(wt, params, lane, 3, ha, hb);
        v4f z = output_layer(wt, params, lane, hb);
        if (lane < 16 && valid) {
            if (RENDER) {
                float4* L = reinterpret_cast<float4*>(out);
                float4 l = L[slot];
                // exp(min(z, 20)): a diverged cache cannot put inf into the film
                L[slot] = make_float4(l.x + thr[0] * expf(fminf(z[0], 20.0f)), l.y + thr[1] * expf(fminf(z[1], 20.0f)),
                                      l.z + thr[2] * expf(fminf(z[2], 20.0f)), l.w);
            } else {
                out[size_t(qi) * 3 + 0] = expf(z[0]);
                out[size_t(qi) * 3 + 1] = expf(z[1]);
                out[size_t(qi) * 3 + 2] = expf(z[2]);
            }
        }
    }
}

__device__ __forceinline__ size_t quad_index(uint32_t s, int f, int F) {
    return ((size_t(s >> 2) * F + f) << 2) + (s & 3u);
}

// Forward + backward of 16 samples per wave.  Relative L2 loss
// (y - t)^2 / (sg(y)^2 + eps) averaged over samples and channels.
// Samples of slice j of `steps` interleaved slices of a batch whose size is
// min(*n_dev, n_host) (n_host alone without a device count): samples j,
// j + steps, j + 2 steps, ...
__device__ __forceinline__ uint32_t slice_count(uint32_t n_host, const uint32_t* n_dev, uint32_t steps, uint32_t j) {
    const uint32_t total = n_dev ? min(*n_dev, n_host) : n_host;
    return total > j ? (total - j + steps - 1) / steps : 0u;
}

__global__ __launch_bounds__(kThreads) void k_cache_fwd_bwd(const v4s* __restrict__ wt_g, const v4s* __restrict__ wtb_g,
                                                            const float* __restrict__ params, Bounds B,
                                                            const float* __restrict__ queries,
                                                            const float* __restrict__ targets, uint32_t n_host,
                                                            const uint32_t* __restrict__ n_dev, uint32_t stride,
                                                            uint32_t offset, float loss_eps, short* __restrict__ act,
                                                            short* __restrict__ dz, float* __restrict__ loss_out) {
    const uint32_t n = slice_count(n_host, n_dev, stride, offset), n_pad = (n + 15) / 16 * 16;
    if (blockIdx.x * (kThreads / 64) * 16 >= n_pad) return;  // no sample group for this block (uniform)
    __shared__ v4s wt[kFwdTiles * 64];
    __shared__ v4s wtb[kBwdTiles * 64];
    load_tiles(wt, wt_g, kFwdTiles * 64);
    load_tiles(wtb, wtb_g, kBwdTiles * 64);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const uint32_t wave = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    const uint32_t nwaves = gridDim.x * (kThreads / 64);
    const float norm = 1.0f / (float(n) * kOut);
    float loss = 0.0f;
    for (uint32_t g = wave; g * 16 < n_pad; g += nwaves) {
        const uint32_t s = g * 16 + (lane & 15);
        const bool valid = s < n;
        float q[12];
        const size_t row = size_t(s) * stride + offset;  // sample s of the slice
#pragma unroll
        for (int k = 0; k < 12; ++k) q[k] = valid ? queries[row * 12 + k] : 0.0f;
        float p01[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) p01[a] = (q[a] - B.lo[a]) * B.inv[a];
        v4s x[6];
        v4s h[4][4];
        layer0(wt, params, lane, p01, q, valid, h[0], x);
        hidden(wt, params, lane, 1, h[0], h[1]);
        hidden(wt, params, lane, 2, h[1], h[2]);
        hidden(wt, params, lane, 3, h[2], h[3]);
        v4f z = output_layer(wt, params, lane, h[3]);
        // output gradient (rows 0..2 live on lanes 0..15)
        v4f d4 = {0.0f, 0.0f, 0.0f, 0.0f};
        if (lane < 16 && valid) {
#pragma unroll
            for (int e = 0; e < kOut; ++e) {
                float y = expf(z[e]);
                float r = y - targets[row * 3 + e];
                float den = y * y + loss_eps;
                loss += r * r / den * norm;
                d4[e] = 2.0f * r / den * norm * y;
            }
        }
        // store layer inputs (x, h1..h4) and dz4
        const int row0 = 4 * (lane >> 4);
#pragma unroll
        for (int kt = 0; kt < 6; ++kt)
#pragma unroll
            for (int e = 0; e < 4; ++e) act[quad_index(s, kActBase[0] + 16 * kt + row0 + e, kActF)] = x[kt][e];
#pragma unroll
        for (int l = 0; l < 4; ++l)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                for (int e = 0; e < 4; ++e) act[quad_index(s, kActBase[l + 1] + 16 * nt + row0 + e, kActF)] = h[l][nt][e];
        v4s dcur[4];
        dcur[0] = pack4(d4);
#pragma unroll
        for (int e = 0; e < 4; ++e) dz[quad_index(s, kDzBase[4] + row0 + e, kDzF)] = dcur[0][e];
        // layer 4 -> dz3 (masked by h4 > 0)
        v4s dn[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mfma(wtb[bwd_tile(4, nt, 0) * 64 + lane], dcur[0], acc);
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[e] = bf2f(h[3][nt][e]) > 0.0f ? acc[e] : 0.0f;
            dn[nt] = pack4(acc);
#pragma unroll
            for (int e = 0; e < 4; ++e) dz[quad_index(s, kDzBase[3] + 16 * nt + row0 + e, kDzF)] = dn[nt][e];
        }
        // layers 3, 2, 1 -> dz2, dz1, dz0
#pragma unroll
        for (int l = 3; l >= 1; --l) {
#pragma unroll
            for (int k = 0; k < 4; ++k) dcur[k] = dn[k];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int kt = 0; kt < 4; kt += 2) {
                    __builtin_amdgcn_sched_barrier(0);
                    acc = mfma2(wtb[bwd_tile(l, nt, kt) * 64 + lane], wtb[bwd_tile(l, nt, kt + 1) * 64 + lane],
                                dcur[kt], dcur[kt + 1], acc);
                }
#pragma unroll
                for (int e = 0; e < 4; ++e) acc[e] = bf2f(h[l - 1][nt][e]) > 0.0f ? acc[e] : 0.0f;
                dn[nt] = pack4(acc);
#pragma unroll
                for (int e = 0; e < 4; ++e) dz[quad_index(s, kDzBase[l - 1] + 16 * nt + row0 + e, kDzF)] = dn[nt][e];
            }
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o, 64);
    if (lane == 0 && loss != 0.0f) atomicAdd(loss_out, loss);
}

struct WgTile {
    int l, ot, it;
};
__device__ __forceinline__ WgTile wg_tile(int t) {
    int l = 0;
    while (l < kLayers - 1 && t >= kFwdBase[l + 1]) ++l;
    int r = t - kFwdBase[l];
    return {l, r / ktiles(l), r % ktiles(l)};
}

// dW_l tile (ot, it) = sum over samples of dz_l^T a_l, one split of the
// samples per wave: A = dz tile (rows = out features, k = samples), B = a tile
// (k = samples, columns = in features), both read as 4 consecutive samples.
// The waves of input tile 0 also sum their split's dz rows: the bias
// gradient partials (partial_b[split][dz feature]).
__global__ __launch_bounds__(kThreads) void k_cache_wgrad(const short* __restrict__ act, const short* __restrict__ dz,
                                                          uint32_t n_host, const uint32_t* __restrict__ n_dev,
                                                          uint32_t stride, uint32_t offset, float* __restrict__ partial,
                                                          float* __restrict__ partial_b) {
    const uint32_t n_pad = (slice_count(n_host, n_dev, stride, offset) + 15) / 16 * 16;
    const int lane = threadIdx.x & 63;
    const uint32_t wave = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    if (wave >= uint32_t(kWgTiles * kSplits)) return;
    const int tile = int(wave % kWgTiles), split = int(wave / kWgTiles);
    WgTile T = wg_tile(tile);
    const uint32_t steps = n_pad / 16;
    const uint32_t per = (steps + kSplits - 1) / kSplits;
    const uint32_t k0 = split * per, k1 = min(steps, k0 + per);
    const int fo = kDzBase[T.l] + 16 * T.ot + (lane & 15);
    const int fi = kActBase[T.l] + 16 * T.it + (lane & 15);
    v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;
    uint32_t k = k0;
    for (; k + 1 < k1; k += 2) {  // two 16-sample steps per 16x16x32 product
        const uint32_t q0 = k * 4 + (lane >> 4), q1 = q0 + 4;  // samples 4 q .. 4 q + 3
        const v4s a0 = *reinterpret_cast<const v4s*>(dz + ((size_t(q0) * kDzF + fo) << 2));
        const v4s b0 = *reinterpret_cast<const v4s*>(act + ((size_t(q0) * kActF + fi) << 2));
        const v4s a1 = *reinterpret_cast<const v4s*>(dz + ((size_t(q1) * kDzF + fo) << 2));
        const v4s b1 = *reinterpret_cast<const v4s*>(act + ((size_t(q1) * kActF + fi) << 2));
        acc = mfma2(a0, a1, b0, b1, acc);
        bsum += (bf2f(a0[0]) + bf2f(a0[1])) + (bf2f(a0[2]) + bf2f(a0[3]));
        bsum += (bf2f(a1[0]) + bf2f(a1[1])) + (bf2f(a1[2]) + bf2f(a1[3]));
    }
    if (k < k1) {
        const uint32_t quad = k * 4 + (lane >> 4);
        const v4s a = *reinterpret_cast<const v4s*>(dz + ((size_t(quad) * kDzF + fo) << 2));
        const v4s b = *reinterpret_cast<const v4s*>(act + ((size_t(quad) * kActF + fi) << 2));
        acc = mfma(a, b, acc);
        bsum += (bf2f(a[0]) + bf2f(a[1])) + (bf2f(a[2]) + bf2f(a[3]));
    }
    float4* out = reinterpret_cast<float4*>(partial + (size_t(split) * kWgTiles + tile) * 256);
    out[lane] = make_float4(acc[0], acc[1], acc[2], acc[3]);
    if (T.it == 0) {
        bsum += __shfl_xor(bsum, 16, 64);
        bsum += __shfl_xor(bsum, 32, 64);
        if (lane < 16) partial_b[size_t(split) * kDzF + fo] = bsum;
    }
}

// Sums the split partials in a fixed order into the FP32 gradient (oracle
// layout): weight tiles and bias sums.
__global__ __launch_bounds__(kThreads) void k_cache_reduce(const float* __restrict__ partial,
                                                           const float* __restrict__ partial_b,
                                                           float* __restrict__ grad) {
    const uint32_t p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= uint32_t(kParams)) return;
    int l = 0;
    while (l < kLayers - 1 && int(p) >= kOffW[l + 1]) ++l;
    int r = int(p) - kOffW[l];
    float g = 0.0f;
    if (r < fan_out(l) * fan_in(l)) {
        int o = r / fan_in(l), i = r % fan_in(l);
        int tile = fwd_tile(l, o / 16, i / 16);
        int row = o % 16, col = i % 16;
        int lane = 16 * (row >> 2) + col, e = row & 3;
        for (int s = 0; s < kSplits; ++s) g += partial[(size_t(s) * kWgTiles + tile) * 256 + lane * 4 + e];
    } else {
        int f = kDzBase[l] + (r - fan_out(l) * fan_in(l));
        for (int s = 0; s < kSplits; ++s) g += partial_b[size_t(s) * kDzF + f];
    }
    grad[p] = g;
}

// Adam (Kingma & Ba) on the FP32 masters; step t = *step + 1 (bias
// corrections 1 - beta^t).  A slice without samples takes no step.
__global__ void k_cache_adam(float* __restrict__ params, float* __restrict__ m, float* __restrict__ v,
                             const float* __restrict__ grad, float lr, float b1, float b2, float eps, uint32_t n_host,
                             const uint32_t* __restrict__ n_dev, uint32_t stride, uint32_t offset,
                             const uint32_t* __restrict__ step) {
    if (slice_count(n_host, n_dev, stride, offset) == 0) return;
    uint32_t p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= uint32_t(kParams)) return;
    const float t = float(*step + 1u);
    const float c1 = 1.0f - powf(b1, t), c2 = 1.0f - powf(b2, t);
    float g = grad[p];
    float mm = b1 * m[p] + (1.0f - b1) * g;
    float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    params[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// BF16 A tiles from the FP32 masters: forward W tiles and backward W^T tiles.
__global__ void k_cache_pack(const float* __restrict__ params, v4s* __restrict__ fwd, v4s* __restrict__ bwd) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    const uint32_t nf = kFwdTiles * 64, nb = kBwdTiles * 64;
    if (i >= nf + nb) return;
    v4s out;
    if (i < nf) {
        int t = int(i / 64), lane = int(i % 64);
        int l = 0;
        while (l < kLayers - 1 && t >= kFwdBase[l + 1]) ++l;
        int nt = (t - kFwdBase[l]) / ktiles(l), kt = (t - kFwdBase[l]) % ktiles(l);
        int o = 16 * nt + (lane & 15);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int in = 16 * kt + 4 * (lane >> 4) + e;
            out[e] = f2bf(o < fan_out(l) && in < fan_in(l) ? params[kOffW[l] + o * fan_in(l) + in] : 0.0f);
        }
        fwd[i] = out;
    } else {
        uint32_t j = i - nf;
        int t = int(j / 64), lane = int(j % 64);
        int l = 1;
        while (l < kLayers - 1 && t >= kBwdBase[l + 1]) ++l;
        int nt = (t - kBwdBase[l]) / ntiles(l), kt = (t - kBwdBase[l]) % ntiles(l);
        int in = 16 * nt + (lane & 15);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int o = 16 * kt + 4 * (lane >> 4) + e;
            out[e] = f2bf(o < fan_out(l) && in < fan_in(l) ? params[kOffW[l] + o * fan_in(l) + in] : 0.0f);
        }
        bwd[j] = out;
    }
}

// Counts the Adam step just taken (after k_cache_adam, which read the count).
__global__ void k_cache_step_done(uint32_t n_host, const uint32_t* __restrict__ n_dev, uint32_t stride, uint32_t offset,
                                  uint32_t* __restrict__ step) {
    if (slice_count(n_host, n_dev, stride, offset) > 0) *step += 1u;
}

// End of a rendered frame: the training-sample count is frozen for the
// optimizer steps (kCcLastSamples; the next frame clears kCcSamples) and, when
// adaptive, the training-path stride is steered towards 0.9 max_batch
// samples per frame (kCcStride).
__global__ void k_cache_frame_end(uint32_t* __restrict__ ctr, uint32_t max_batch, uint32_t adaptive) {
    const uint32_t got = ctr[kCcSamples];
    ctr[kCcLastSamples] = min(got, max_batch);
    if (adaptive) {
        const float cur = float(ctr[kCcStride]);
        const float next = got ? cur * float(got) / (0.9f * float(max_batch)) : cur * 0.5f;
        ctr[kCcStride] = uint32_t(fminf(fmaxf(next, 1.0f), float(1u << 20)));
    }
}

__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }

__global__ void k_zero_f(float* p, uint32_t n) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

__global__ void k_zero_u(uint32_t* p, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

}  // namespace

// Device counters of a cache: kCacheCounters words laid out as kCc* says.

struct rc_cache {
    rc_ctx* ctx = nullptr;
    rc_cache_cfg cfg{};
    Bounds bounds{};
    float* params = nullptr;  // FP32 masters
    float* m = nullptr;
    float* v = nullptr;
    float* grad = nullptr;
    v4s* fwd = nullptr;
    v4s* bwd = nullptr;
    float* loss = nullptr;
    uint32_t* counters = nullptr;  // kCacheCounters words, read and written by the kernels
    bool stride_set = false;       // counters[kCcStride] holds an adaptive stride
    DevBuf act, dz, partial, queries, targets, out;
    DevBuf pass_q, pass_r;  // queries and training records of one render pass
    // The optimizer steps after a rendered frame run on their own stream,
    // while the host goes on and the next frame traverses; the next cache
    // inference (or any call touching the weights) waits for ev_trained.
    hipStream_t train_stream = nullptr;
    hipEvent_t ev_frame = nullptr, ev_t0 = nullptr, ev_trained = nullptr;
    bool pending = false;     // steps queued since the last wait
    uint32_t last_steps = 0;  // steps of the last frame (for its mean loss)
};

namespace {

int cache_pack(rc_ctx* c, rc_cache* k) {
    const uint32_t n = (kFwdTiles + kBwdTiles) * 64;
    k_cache_pack<<<(n + kThreads - 1) / kThreads, kThreads, 0, c->stream>>>(k->params, k->fwd, k->bwd);
    RC_TRY_HIP(hipGetLastError());
    return RC_OK;
}

uint32_t waves_grid(uint32_t n_pad) {
    uint32_t groups = n_pad / 16;
    return std::max(1u, std::min(2048u, (groups + 3) / 4));
}

// Makes stream st wait for the optimizer steps in flight (GPU-side wait).
int wait_training(rc_cache* k, hipStream_t st) {
    if (k->pending) RC_TRY_HIP(hipStreamWaitEvent(st, k->ev_trained, 0));
    return RC_OK;
}

// Gradient of the loss over slice j of `steps` slices of the batch in
// (d_queries, d_targets), whose size is min(*n_dev, n_host); on stream st.
int grad_device(rc_ctx* c, hipStream_t st, rc_cache* k, const float* d_queries, const float* d_targets,
                uint32_t n_host, const uint32_t* n_dev, uint32_t steps, uint32_t j) {
    const uint32_t n_max = (n_host - std::min(n_host, j) + steps - 1) / steps, n_pad = (n_max + 15) / 16 * 16;
    RC_TRY(k->act.ensure(size_t(std::max(n_pad, 16u)) * kActF * sizeof(short)));
    RC_TRY(k->dz.ensure(size_t(std::max(n_pad, 16u)) * kDzF * sizeof(short)));
    RC_TRY(k->partial.ensure(size_t(kSplits) * (kWgTiles * 256 + kDzF) * sizeof(float)));
    float* partial_b = k->partial.as<float>() + size_t(kSplits) * kWgTiles * 256;
    k_cache_fwd_bwd<<<waves_grid(std::max(n_pad, 16u)), kThreads, 0, st>>>(
        k->fwd, k->bwd, k->params, k->bounds, d_queries, d_targets, n_host, n_dev, steps, j, k->cfg.loss_eps,
        k->act.as<short>(), k->dz.as<short>(), k->loss);
    const uint32_t wg_waves = kWgTiles * kSplits;
    k_cache_wgrad<<<(wg_waves + 3) / 4, kThreads, 0, st>>>(k->act.as<short>(), k->dz.as<short>(), n_host, n_dev,
                                                          steps, j, k->partial.as<float>(), partial_b);
    k_cache_reduce<<<(kParams + kThreads - 1) / kThreads, kThreads, 0, st>>>(k->partial.as<float>(), partial_b,
                                                                             k->grad);
    RC_TRY_HIP(hipGetLastError());
    return RC_OK;
}

// One Adam step on the gradient of slice j (skipped when it has no samples).
int adam_device(rc_ctx* c, hipStream_t st, rc_cache* k, uint32_t n_host, const uint32_t* n_dev, uint32_t steps,
                uint32_t j) {
    k_cache_adam<<<(kParams + kThreads - 1) / kThreads, kThreads, 0, st>>>(
        k->params, k->m, k->v, k->grad, k->cfg.lr, k->cfg.beta1, k->cfg.beta2, k->cfg.adam_eps, n_host, n_dev, steps, j,
        k->counters + kCcSteps);
    k_cache_step_done<<<1, 1, 0, st>>>(n_host, n_dev, steps, j, k->counters + kCcSteps);
    const uint32_t n = (kFwdTiles + kBwdTiles) * 64;
    k_cache_pack<<<(n + kThreads - 1) / kThreads, kThreads, 0, st>>>(k->params, k->fwd, k->bwd);
    RC_TRY_HIP(hipGetLastError());
    return RC_OK;
}

}  // namespace

// Device-side entry points used by the renderer.
int rc_cache_apply_device(rc_ctx* c, hipStream_t st, rc_cache* k, const float4* queries, const uint32_t* n_dev,
                          uint32_t n_bound, float4* L) {
    if (n_bound == 0) return RC_OK;
    RC_TRY(wait_training(k, st));
    uint32_t grid = waves_grid((n_bound + 15) / 16 * 16);
    k_cache_infer<true><<<grid, kThreads, 0, st>>>(k->fwd, k->params, k->bounds, reinterpret_cast<const float*>(queries),
                                                   n_bound, n_dev, reinterpret_cast<float*>(L));
    RC_TRY_HIP(hipGetLastError());
    return RC_OK;
}

int rc_cache_frame_io(rc_ctx* c, hipStream_t st, rc_cache* k, uint32_t cache_vertex, size_t pass_paths,
                      uint64_t frame_samples, CacheIO* io) {
    const size_t need_q = pass_paths * 4 * sizeof(float4), need_r = pass_paths * 5 * sizeof(float4);
    const size_t need_bq = size_t(k->cfg.max_batch) * 12 * 4, need_bt = size_t(k->cfg.max_batch) * 3 * 4;
    if (k->pending && (k->pass_q.bytes < need_q || k->pass_r.bytes < need_r || k->queries.bytes < need_bq ||
                       k->targets.bytes < need_bt))
        RC_TRY_HIP(hipEventSynchronize(k->ev_trained));  // buffers the steps in flight read are about to move
    RC_TRY(k->pass_q.ensure(need_q));
    RC_TRY(k->pass_r.ensure(need_r));
    RC_TRY(k->queries.ensure(need_bq));
    RC_TRY(k->targets.ensure(need_bt));
    io->queries = k->pass_q.as<float4>();
    io->records = k->pass_r.as<float4>();
    io->counters = k->counters;
    io->batch_q = k->queries.as<float>();
    io->batch_t = k->targets.as<float>();
    io->batch_cap = k->cfg.max_batch;
    io->cache_vertex = cache_vertex;
    io->train_stride = k->cfg.train_stride;
    io->stride_dev = k->counters + kCcStride;
    io->cap = uint32_t(pass_paths);
    k_zero_u<<<1, 64, 0, st>>>(k->counters, kCcFrameQueries + 1);  // the pass and frame counts
    if (k->cfg.train_stride == 0 && !k->stride_set) {
        // adaptive: the first frame as if every training path reached the
        // cache vertex; later frames steer the count (k_cache_frame_end)
        const uint32_t first =
            uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(frame_samples / k->cfg.max_batch, 1u << 20)));
        k_set_u32<<<1, 1, 0, st>>>(k->counters + kCcStride, first);
        k->stride_set = true;
    }
    RC_TRY_HIP(hipGetLastError());
    return RC_OK;
}

rc_ctx* rc_cache_owner(const rc_cache* k) { return k->ctx; }

int rc_cache_frame_end(rc_ctx* c, hipStream_t st, rc_cache* k, uint32_t steps, uint32_t counts[2]) {
    k_cache_frame_end<<<1, 1, 0, st>>>(k->counters, k->cfg.max_batch, k->cfg.train_stride == 0 ? 1u : 0u);
    uint32_t h[2];  // training samples, cache queries of the frame
    RC_TRY_HIP(hipMemcpyAsync(h, k->counters + kCcSamples, sizeof(h), hipMemcpyDeviceToHost, st));
    k->last_steps = steps;
    if (steps > 0) {
        // the optimizer steps wait for the frame's samples on the cache's stream
        RC_TRY_HIP(hipEventRecord(k->ev_frame, st));
        RC_TRY_HIP(hipStreamWaitEvent(k->train_stream, k->ev_frame, 0));
        RC_TRY_HIP(hipEventRecord(k->ev_t0, k->train_stream));
        k_zero_f<<<1, kThreads, 0, k->train_stream>>>(k->loss, 1);
        for (uint32_t j = 0; j < steps; ++j) {
            RC_TRY(grad_device(c, k->train_stream, k, k->queries.as<float>(), k->targets.as<float>(), k->cfg.max_batch,
                               k->counters + kCcLastSamples, steps, j));
            RC_TRY(adam_device(c, k->train_stream, k, k->cfg.max_batch, k->counters + kCcLastSamples, steps, j));
        }
        RC_TRY_HIP(hipEventRecord(k->ev_trained, k->train_stream));
        k->pending = true;
        c->train_ev0 = k->ev_t0;
        c->train_ev1 = k->ev_trained;
    }
    RC_TRY_HIP(hipStreamSynchronize(st));
    counts[0] = h[1];
    counts[1] = std::min(h[0], k->cfg.max_batch);
    return RC_OK;
}

int rc_cache_last_loss(rc_ctx* c, rc_cache* k, float* loss) {
    *loss = 0.0f;
    if (k->last_steps == 0) return RC_OK;
    RC_TRY_HIP(hipEventSynchronize(k->ev_trained));
    float l = 0.0f;
    RC_TRY_HIP(hipMemcpy(&l, k->loss, 4, hipMemcpyDeviceToHost));
    *loss = l / float(k->last_steps);
    return RC_OK;
}

extern "C" int rc_cache_create(rc_ctx* c, const rc_cache_cfg* cfg, const float bmin[3], const float bmax[3],
                               rc_cache** out) {
    RC_TRY(rc_check_ctx(c, "rc_cache_create"));
    if (!cfg || !bmin || !bmax || !out) return rc_host_fail(RC_ERR_INVALID, "rc_cache_create: null argument");
    if (cfg->max_batch == 0 || cfg->max_batch > (1u << 24) || !(cfg->lr > 0.0f))
        return rc_host_fail(RC_ERR_INVALID, "rc_cache_create: bad configuration");
    RC_TRY_HIP(hipSetDevice(c->device));
    rc_cache* k = new (std::nothrow) rc_cache();
    if (!k) return rc_host_fail(RC_ERR_ALLOC, "rc_cache_create: out of memory");
    k->ctx = c;
    k->cfg = *cfg;
    for (int a = 0; a < 3; ++a) {
        float ext = bmax[a] - bmin[a];
        k->bounds.lo[a] = bmin[a];
        k->bounds.inv[a] = ext > 0.0f ? 1.0f / ext : 1.0f;
    }
    auto fail = [&](const char* what) {
        rc_cache_free(k);
        return rc_host_fail(RC_ERR_ALLOC, std::string("rc_cache_create: ") + what);
    };
    if (hipMalloc(&k->params, kParams * 4) != hipSuccess || hipMalloc(&k->m, kParams * 4) != hipSuccess ||
        hipMalloc(&k->v, kParams * 4) != hipSuccess || hipMalloc(&k->grad, kParams * 4) != hipSuccess ||
        hipMalloc(&k->fwd, kFwdTiles * 64 * sizeof(v4s)) != hipSuccess ||
        hipMalloc(&k->bwd, kBwdTiles * 64 * sizeof(v4s)) != hipSuccess || hipMalloc(&k->loss, 4) != hipSuccess ||
        hipMalloc(&k->counters, kCacheCounters * sizeof(uint32_t)) != hipSuccess)
        return fail("device allocation failed");
    if (hipStreamCreateWithFlags(&k->train_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&k->ev_frame, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&k->ev_t0) != hipSuccess || hipEventCreate(&k->ev_trained) != hipSuccess)
        return fail("stream or event creation failed");
    k_zero_u<<<1, 64, 0, c->stream>>>(k->counters, kCacheCounters);
    // He-uniform hidden layers, small output layer with the output starting near 0.5
    std::vector<float> p(kParams, 0.0f);
    std::mt19937 rng(cfg->seed);
    for (int l = 0; l < kLayers; ++l) {
        float lim = l == 4 ? 0.01f : std::sqrt(6.0f / float(fan_in(l)));
        std::uniform_real_distribution<float> U(-lim, lim);
        for (int i = 0; i < fan_out(l) * fan_in(l); ++i) p[kOffW[l] + i] = U(rng);
        for (int o = 0; o < fan_out(l); ++o) p[kOffB[l] + o] = l == 4 ? std::log(0.5f) : 0.0f;
    }
    int rc = rc_cache_set_params(k, p.data());
    if (rc != RC_OK) {
        rc_cache_free(k);
        return rc;
    }
    *out = k;
    return RC_OK;
}

namespace {
// Host-side wait for the optimizer steps in flight (the synchronous API calls
// below read or replace what they write).
int settle(rc_cache* k) {
    if (k->pending) RC_TRY_HIP(hipEventSynchronize(k->ev_trained));
    k->pending = false;
    return RC_OK;
}
}  // namespace

extern "C" int rc_cache_free(rc_cache* k) {
    if (!k) return RC_OK;
    (void)hipStreamSynchronize(k->ctx->stream);
    if (k->train_stream) (void)hipStreamSynchronize(k->train_stream);
    if (k->ctx->train_ev1 == k->ev_trained) k->ctx->train_ev0 = k->ctx->train_ev1 = nullptr;
    if (k->ev_frame) (void)hipEventDestroy(k->ev_frame);
    if (k->ev_t0) (void)hipEventDestroy(k->ev_t0);
    if (k->ev_trained) (void)hipEventDestroy(k->ev_trained);
    if (k->train_stream) (void)hipStreamDestroy(k->train_stream);
    float* fp[] = {k->params, k->m, k->v, k->grad, k->loss};
    for (float* q : fp)
        if (q) (void)hipFree(q);
    if (k->fwd) (void)hipFree(k->fwd);
    if (k->bwd) (void)hipFree(k->bwd);
    if (k->counters) (void)hipFree(k->counters);
    DevBuf* b[] = {&k->act, &k->dz, &k->partial, &k->queries, &k->targets, &k->out, &k->pass_q, &k->pass_r};
    for (DevBuf* x : b) x->release();
    delete k;
    return RC_OK;
}

extern "C" int rc_cache_set_params(rc_cache* k, const float* params) {
    if (!k || !params) return rc_host_fail(RC_ERR_INVALID, "rc_cache_set_params: null argument");
    rc_ctx* c = k->ctx;
    RC_TRY(settle(k));
    RC_TRY_HIP(hipMemcpyAsync(k->params, params, kParams * 4, hipMemcpyHostToDevice, c->stream));
    k_zero_f<<<(kParams + kThreads - 1) / kThreads, kThreads, 0, c->stream>>>(k->m, kParams);
    k_zero_f<<<(kParams + kThreads - 1) / kThreads, kThreads, 0, c->stream>>>(k->v, kParams);
    k_set_u32<<<1, 1, 0, c->stream>>>(k->counters + kCcSteps, 0u);  // Adam step count
    RC_TRY(cache_pack(c, k));
    RC_TRY_HIP(hipStreamSynchronize(c->stream));
    return RC_OK;
}

extern "C" int rc_cache_get_params(rc_cache* k, float* params) {
    if (!k || !params) return rc_host_fail(RC_ERR_INVALID, "rc_cache_get_params: null argument");
    RC_TRY(settle(k));
    RC_TRY_HIP(hipMemcpyAsync(params, k->params, kParams * 4, hipMemcpyDeviceToHost, k->ctx->stream));
    RC_TRY_HIP(hipStreamSynchronize(k->ctx->stream));
    return RC_OK;
}

namespace {
int stage(rc_ctx* c, rc_cache* k, const float* q, const float* t, size_t n) {
    RC_TRY(k->queries.ensure(n * 12 * 4));
    RC_TRY_HIP(hipMemcpyAsync(k->queries.p, q, n * 12 * 4, hipMemcpyHostToDevice, c->stream));
    if (t) {
        RC_TRY(k->targets.ensure(n * 3 * 4));
        RC_TRY_HIP(hipMemcpyAsync(k->targets.p, t, n * 3 * 4, hipMemcpyHostToDevice, c->stream));
    }
    return RC_OK;
}
}  // namespace

extern "C" int rc_cache_infer(rc_ctx* c, rc_cache* k, const float* q, size_t n, float* out) {
    RC_TRY(rc_check_ctx(c, "rc_cache_infer"));
    if (!k || k->ctx != c || (n && (!q || !out))) return rc_host_fail(RC_ERR_INVALID, "rc_cache_infer: bad argument");
    if (n == 0) return RC_OK;
    if (n > (1u << 28)) return rc_host_fail(RC_ERR_INVALID, "rc_cache_infer: too many queries");
    RC_TRY(settle(k));
    RC_TRY(stage(c, k, q, nullptr, n));
    RC_TRY(k->out.ensure(n * 3 * 4));
    uint32_t n_pad = uint32_t((n + 15) / 16 * 16);
    k_cache_infer<false><<<waves_grid(n_pad), kThreads, 0, c->stream>>>(k->fwd, k->params, k->bounds,
                                                                        k->queries.as<float>(), uint32_t(n), nullptr,
                                                                        k->out.as<float>());
    RC_TRY_HIP(hipGetLastError());
    RC_TRY_HIP(hipMemcpyAsync(out, k->out.p, n * 3 * 4, hipMemcpyDeviceToHost, c->stream));
    RC_TRY_HIP(hipStreamSynchronize(c->stream));
    return RC_OK;
}

extern "C" int rc_cache_gradient(rc_ctx* c, rc_cache* k, const float* q, const float* t, size_t n, float* grad,
                                 float* loss) {
    RC_TRY(rc_check_ctx(c, "rc_cache_gradient"));
    if (!k || k->ctx != c || !q || !t || n == 0) return rc_host_fail(RC_ERR_INVALID, "rc_cache_gradient: bad argument");
    if (n > k->cfg.max_batch) return rc_host_fail(RC_ERR_INVALID, "rc_cache_gradient: batch larger than max_batch");
    RC_TRY(settle(k));
    RC_TRY(stage(c, k, q, t, n));
    k_zero_f<<<1, kThreads, 0, c->stream>>>(k->loss, 1);
    RC_TRY(grad_device(c, c->stream, k, k->queries.as<float>(), k->targets.as<float>(), uint32_t(n), nullptr, 1, 0));
    if (grad) RC_TRY_HIP(hipMemcpyAsync(grad, k->grad, kParams * 4, hipMemcpyDeviceToHost, c->stream));
    float l = 0.0f;
    RC_TRY_HIP(hipMemcpyAsync(&l, k->loss, 4, hipMemcpyDeviceToHost, c->stream));
    RC_TRY_HIP(hipStreamSynchronize(c->stream));
    if (loss) *loss = l;
    return RC_OK;
}

extern "C" int rc_cache_train_step(rc_ctx* c, rc_cache* k, const float* q, const float* t, size_t n, float* loss) {
    RC_TRY(rc_cache_gradient(c, k, q, t, n, nullptr, loss));
    RC_TRY(adam_device(c, c->stream, k, uint32_t(n), nullptr, 1, 0));
    RC_TRY_HIP(hipStreamSynchronize(c->stream));
    return RC_OK;
}
