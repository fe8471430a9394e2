// cache.hip — neural radiance cache on gfx950: multiresolution hash-grid encoding, a 6 x 64x64
// MLP with BF16 MFMA forward and backward (FP32 accumulation), relative L2 loss, Adam.
//
// No reference counterpart (the reference has no learned component). Published methods:
// hash encoding (Mueller et al., Instant NGP, 2022), neural radiance caching and its relative
// L2 loss (Mueller et al., 2021), Adam (Kingma & Ba, 2015).
//
// MFMA: v_mfma_f32_16x16x32_bf16, D[i][j] = sum_k A[i][k] B[k][j] + C[i][j] with the wave64
// operand layout  A: lane = i + 16*(k/8), element k%8;  B: lane = j + 16*(k/8), element k%8;
// D: lane = j + 16*(i/4), register i%4.
// v_mfma_f32_32x32x16_bf16: A: lane = i + 32*(k/8), element k%8;  B: lane = j + 32*(k/8),
// element k%8;  D: lane = j + 32*((i/4)%2), register 4*(i/8) + i%4.
// Inference forward (k_mlp_infer, one wave per 64 queries) uses the 32x32x16 form with the same
// idea as below (the permutation phi32); the training forward keeps the 16x16x32 form, whose
// D fragments are also written out as the activations the backward kernels read.
// Training forward (fused, one wave per 16 queries, weights in LDS): rows = output neurons, columns =
// queries. The D fragment of a layer (neurons 16t + 4g + r of query j in lane j + 16g) is used
// directly as the next layer's B operand: the B element k = 32b + 8g + e of lane j + 16g is
// neuron phi(k) = 32b + 4g + e for e < 4 and 32b + 16 + 4g + (e - 4) for e >= 4, so the next
// layer's weights are packed with their input index permuted by phi and activations never
// leave registers. The first layer reads its input from memory in natural order.
// Backward: one MFMA kernel chains the data gradients through all six layers (weights in LDS,
// each layer's D fragment fed back as the next B operand, as in the forward), one MFMA kernel
// forms all six weight gradients, then an atomic scatter of the hash-grid gradient and Adam.
#include "internal.h"

#include <cstring>
#include <new>
#include <random>

namespace {

constexpr int kW = 64;         // layer width
constexpr int kLayers = 6;     // 64x64 layers; ReLU after the first five
constexpr uint32_t kMaxLossBlocks = kLossCountSlot;  // blocks of k_mlp_train (one loss partial each)
constexpr int kEnc = 32;       // hash-grid features (16 levels x 2)
constexpr int kMaxLevels = 16;
constexpr int kFragBytes = 16;                       // 8 bf16 per lane per fragment
constexpr int kLayerPack = 4 * 2 * 64 * kFragBytes;  // 4 row tiles x 2 k-blocks x 64 lanes = 8 KB

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef short s16x2 __attribute__((ext_vector_type(2)));

__host__ __device__ inline uint32_t bf16_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) {
        if (u & 0x007FFFFFu) u |= 0x00400000u;
        return u >> 16;
    }
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}
__host__ __device__ inline float bf16_float(uint32_t b) {
    const uint32_t u = b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
__device__ inline uint32_t pack2(float lo, float hi) { return bf16_bits(lo) | (bf16_bits(hi) << 16); }
__device__ inline bf16x8 as_frag(uint4 v) { return __builtin_bit_cast(bf16x8, v); }
__device__ inline f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
__device__ inline f32x16 mfma32(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
// relu(lo), relu(hi) as two BF16 (round to nearest even, v_cvt_pk_bf16_f32), the ReLU applied to
// the BF16 bit patterns as signed 16-bit integers (negative values and -0 become +0); equal to
// pack2(fmaxf(lo, 0), fmaxf(hi, 0)) for every input that is not a NaN (FP32 denormals kept).
__device__ inline uint32_t relu_pack2(float lo, float hi) {
    const bf16x2 b = __builtin_convertvector((f32x2){lo, hi}, bf16x2);
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s16x2, b), (s16x2){0, 0}));
}

// k-slot of the 32x32x16 B operand -> neuron of the previous layer's D fragments: slot
// k = 16s + 8h + e of lane j + 32h holds register 8(s&1) + e of row tile s>>1.
__host__ __device__ inline int phi32(int k) {
    const int s = k >> 4, h = (k >> 3) & 1, e = k & 7;
    return 32 * (s >> 1) + 16 * (s & 1) + 8 * (e >> 2) + 4 * h + (e & 3);
}

// k-slot of the B operand -> neuron of the previous layer's D fragment
__host__ __device__ inline int phi(int k) {
    const int b = k >> 5, g = (k >> 3) & 3, e = k & 7;
    return 32 * b + (e < 4 ? 4 * g + e : 16 + 4 * g + (e - 4));
}

struct GridDev {
    uint32_t levels, log2_table;
    float res[kMaxLevels];
    uint32_t dense[kMaxLevels];   // 1 if (res + 1)^3 <= table size
    float lo[3], inv_ext[3];
};

__device__ __forceinline__ void corners(const GridDev& g, uint32_t l, const float p[3], uint32_t idx[8], float w[8]) {
    const float res_f = g.res[l];
    const uint32_t res = (uint32_t)res_f;
    const uint32_t tsize = 1u << g.log2_table;
    float fr[3];
    uint32_t c0[3];
    for (int k = 0; k < 3; ++k) {
        float u = (p[k] - g.lo[k]) * g.inv_ext[k];
        u = fminf(fmaxf(u, 0.0f), 1.0f);
        const float x = u * res_f;
        float fl = floorf(x);
        if (fl >= res_f) fl = res_f - 1.0f;
        c0[k] = (uint32_t)fl;
        fr[k] = x - fl;
    }
    const uint32_t side = res + 1u;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const uint32_t dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
        const float wx = dx ? fr[0] : 1.0f - fr[0];
        const float wy = dy ? fr[1] : 1.0f - fr[1];
        const float wz = dz ? fr[2] : 1.0f - fr[2];
        w[c] = (wx * wy) * wz;
        const uint32_t x = c0[0] + dx, y = c0[1] + dy, z = c0[2] + dz;
        idx[c] = g.dense[l] ? (x + side * (y + side * z)) : ((x * 1u ^ y * 2654435761u ^ z * 805459861u) & (tsize - 1u));
    }
}

// ---- encoding ------------------------------------------------------------------------------
// Sample counts: n = min(*count, nmax) when count is given (device-driven batches), else nmax;
// rows n..pad32(n)-1 are written as zeros so every 16- and 32-query tile is complete.
// Transposed copies ([feature][query]) use the leading dimension ld >= pad32(nmax).
__device__ __forceinline__ uint32_t batch_n(const uint32_t* count, uint32_t nmax) {
    return count ? min(*count, nmax) : nmax;
}
__device__ __forceinline__ uint32_t pad32(uint32_t n) { return (n + 31u) & ~31u; }

// scale multiplies the interpolated features (1 for the trained table; the bias correction of
// an averaged table otherwise).
// One thread per query.
__global__ __launch_bounds__(256) void k_encode(GridDev g, const float* __restrict__ table, float scale,
                                                const mrt_cache_sample* __restrict__ qs, const uint32_t* count,
                                                uint32_t nmax, uint32_t ld, uint16_t* __restrict__ x0,
                                                uint16_t* __restrict__ x0T, float* __restrict__ clear) {
    if (clear && blockIdx.x == 0 && threadIdx.x == 0) *clear = 0.0f;  // the training step's loss sum
    const uint32_t q = blockIdx.x * 256u + threadIdx.x;
    const uint32_t n = batch_n(count, nmax);
    if (q >= pad32(n)) return;
    uint32_t vals[kW];
    if (q < n) {
        const mrt_cache_sample Q = qs[q];
        const size_t tsize = (size_t)1 << g.log2_table;
#pragma unroll
        for (uint32_t l = 0; l < (uint32_t)kMaxLevels; ++l) {
            float f0 = 0.0f, f1 = 0.0f;
            if (l < g.levels) {
                uint32_t idx[8];
                float w[8];
                corners(g, l, Q.pos, idx, w);
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    const float2 t = *reinterpret_cast<const float2*>(table + ((size_t)l * tsize + idx[c]) * 2);
                    f0 = f0 + w[c] * t.x;
                    f1 = f1 + w[c] * t.y;
                }
            }
            vals[2 * l] = bf16_bits(f0 * scale);
            vals[2 * l + 1] = bf16_bits(f1 * scale);
        }
        for (int k = 0; k < 3; ++k) {
            vals[kEnc + k] = bf16_bits(Q.dir[k]);
            vals[kEnc + 3 + k] = bf16_bits(Q.nrm[k]);
            vals[kEnc + 6 + k] = bf16_bits(Q.alb[k]);
        }
        for (int k = kEnc + 9; k < kW; ++k) vals[k] = 0;
    } else {
#pragma unroll
        for (int k = 0; k < kW; ++k) vals[k] = 0;
    }
    uint4* row = reinterpret_cast<uint4*>(x0 + (size_t)q * kW);
#pragma unroll
    for (int c = 0; c < kW / 8; ++c)
        row[c] = make_uint4(vals[8 * c] | (vals[8 * c + 1] << 16), vals[8 * c + 2] | (vals[8 * c + 3] << 16),
                            vals[8 * c + 4] | (vals[8 * c + 5] << 16), vals[8 * c + 6] | (vals[8 * c + 7] << 16));
    if (x0T) {
#pragma unroll
        for (int k = 0; k < kW; ++k) x0T[(size_t)k * ld + q] = (uint16_t)vals[k];
    }
}

// The same encoding with one thread per (query, level), for small batches (training steps),
// where one thread per query leaves most CUs idle: thread t handles query t / 16 and level
// l = t % 16 (levels >= g.levels give zero features), and writes x0 elements 2l, 2l+1 (the
// level's two features) and 32+2l, 33+2l (direction, normal, albedo for elements 32..40, zeros
// after). Large batches run faster with one thread per query (k_encode).
__global__ __launch_bounds__(256) void k_encode_lv(GridDev g, const float* __restrict__ table, float scale,
                                                const mrt_cache_sample* __restrict__ qs, const uint32_t* count,
                                                uint32_t nmax, uint32_t ld, uint16_t* __restrict__ x0,
                                                uint16_t* __restrict__ x0T, float* __restrict__ clear) {
    if (clear && blockIdx.x == 0 && threadIdx.x == 0) *clear = 0.0f;  // the training step's loss sum
    static_assert(kMaxLevels == 16 && kW == 64, "thread layout: 16 levels, 64 inputs");
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const uint32_t q = t >> 4, l = t & 15u;
    const uint32_t n = batch_n(count, nmax);
    if (q >= pad32(n)) return;
    uint32_t v0 = 0, v1 = 0, v2 = 0, v3 = 0;  // BF16 bits of elements 2l, 2l+1, 32+2l, 33+2l
    if (q < n) {
        const float* qf = reinterpret_cast<const float*>(qs + q);  // pos, dir, nrm, alb at 0, 4, 8, 12
        if (l < g.levels) {
            const float pos[3] = {qf[0], qf[1], qf[2]};
            const size_t tsize = (size_t)1 << g.log2_table;
            uint32_t idx[8];
            float w[8];
            corners(g, l, pos, idx, w);
            float f0 = 0.0f, f1 = 0.0f;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const float2 tv = *reinterpret_cast<const float2*>(table + ((size_t)l * tsize + idx[c]) * 2);
                f0 = f0 + w[c] * tv.x;
                f1 = f1 + w[c] * tv.y;
            }
            v0 = bf16_bits(f0 * scale);
            v1 = bf16_bits(f1 * scale);
        }
        const uint32_t e0 = 2 * l, e1 = 2 * l + 1;  // extra feature i: float 4 + 4 (i / 3) + i % 3
        if (e0 < 9) v2 = bf16_bits(qf[4 + 4 * (e0 / 3) + e0 % 3]);
        if (e1 < 9) v3 = bf16_bits(qf[4 + 4 * (e1 / 3) + e1 % 3]);
    }
    uint32_t* row = reinterpret_cast<uint32_t*>(x0 + (size_t)q * kW);
    row[l] = v0 | (v1 << 16);
    row[16 + l] = v2 | (v3 << 16);
    if (x0T) {
        x0T[(size_t)(2 * l) * ld + q] = (uint16_t)v0;
        x0T[(size_t)(2 * l + 1) * ld + q] = (uint16_t)v1;
        x0T[(size_t)(32 + 2 * l) * ld + q] = (uint16_t)v2;
        x0T[(size_t)(33 + 2 * l) * ld + q] = (uint16_t)v3;
    }
}

// Copies N uint4 (a multiple of 256) from global memory into LDS with a 256-thread block: every
// load of a thread is issued before its first LDS store, so the copy costs about one memory
// latency instead of one per iteration (a plain strided loop waits for each load in turn).
template <int N>
__device__ __forceinline__ void fill_lds(uint4* __restrict__ dst, const uint4* __restrict__ src) {
    static_assert(N % 256 == 0, "whole rounds of 256 threads");
    constexpr int kPer = N / 256;
    uint4 v[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) v[k] = src[threadIdx.x + 256 * k];
#pragma unroll
    for (int k = 0; k < kPer; ++k) dst[threadIdx.x + 256 * k] = v[k];
}

// ---- inference forward ---------------------------------------------------------------------
// wq: kLayers packed layers in the 32x32x16 layout, fragment (T, s, lane) at ((L*2 + T)*4 + s)*64
// + lane holding W_L[32T + (lane & 31)][k], k = 16s + 8(lane >> 5) + e for L = 0 and
// phi32(16s + 8(lane >> 5) + e) for L > 0. One wave per 64 queries (two column groups of 32
// sharing every weight fragment read from LDS); v_mfma_f32_32x32x16_bf16 reads half the LDS
// bytes per FLOP of the 16x16x32 form. The output layer computes row tile 0 only (rows 0..2
// are the RGB outputs). No activations are kept: queries only.
__global__ __launch_bounds__(256) void k_mlp_infer(const uint4* __restrict__ wq, const uint16_t* __restrict__ x0,
                                                   const uint32_t* count, uint32_t nmax, float* __restrict__ y) {
    __shared__ uint4 s_w[kLayers * 2 * 4 * 64];
    fill_lds<kLayers * 2 * 4 * 64>(s_w, wq);
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, h = lane >> 5, j = lane & 31u;
    const uint32_t wave = (blockIdx.x * 256u + threadIdx.x) >> 6, nwaves = (gridDim.x * 256u) >> 6;
    const uint32_t n = batch_n(count, nmax);
    const uint32_t rows = pad32(n);  // rows n..rows-1 of x0 are zeros
    for (uint32_t base = wave * 64u; base < rows; base += nwaves * 64u) {
        // Compiler fence: the weight fragments are re-read from LDS every iteration instead of being
        // hoisted into 176 registers (which would allow one wave per SIMD instead of three).
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        const bool two = base + 32u < rows;  // wave-uniform: the second group of 32 exists
        bf16x8 B[2][4];
#pragma unroll
        for (int gq = 0; gq < 2; ++gq) {
            const uint4* xr = reinterpret_cast<const uint4*>(x0 + (size_t)(base + 32u * gq + j) * kW);
#pragma unroll
            for (int s = 0; s < 4; ++s)
                B[gq][s] = as_frag((gq == 0 || two) ? xr[2 * s + h] : make_uint4(0u, 0u, 0u, 0u));
        }
        f32x16 acc[2][2];
#pragma unroll
        for (int L = 0; L < kLayers; ++L) {
            const uint4* wl = s_w + L * 512;
#pragma unroll
            for (int T = 0; T < (L + 1 < kLayers ? 2 : 1); ++T) {
                f32x16 c0 = {}, c1 = {};
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    if (L == 0 && s == 3) continue;  // inputs 48..63 are zero (41 encoded features)
                    const bf16x8 a = as_frag(wl[(T * 4 + s) * 64 + lane]);
                    c0 = mfma32(a, B[0][s], c0);
                    c1 = mfma32(a, B[1][s], c1);
                }
                acc[0][T] = c0;
                acc[1][T] = c1;
            }
            if (L + 1 < kLayers) {
#pragma unroll
                for (int gq = 0; gq < 2; ++gq)
#pragma unroll
                    for (int T = 0; T < 2; ++T) {
                        const f32x16& c = acc[gq][T];
                        B[gq][2 * T] = as_frag(make_uint4(relu_pack2(c[0], c[1]), relu_pack2(c[2], c[3]),
                                                          relu_pack2(c[4], c[5]), relu_pack2(c[6], c[7])));
                        B[gq][2 * T + 1] = as_frag(make_uint4(relu_pack2(c[8], c[9]), relu_pack2(c[10], c[11]),
                                                              relu_pack2(c[12], c[13]), relu_pack2(c[14], c[15])));
                    }
            }
        }
        if (h == 0) {  // registers 0..3 of lanes 0..31 are rows 0..3 of the output tile
#pragma unroll
            for (int gq = 0; gq < 2; ++gq) {
                const uint32_t q = base + 32u * gq + j;
                if (q < n)
                    *reinterpret_cast<float4*>(y + (size_t)q * 4) =
                        make_float4(acc[gq][0][0], acc[gq][0][1], acc[gq][0][2], 0.0f);
            }
        }
    }
}

// ---- training step of the MLP: forward, loss, data gradients ----------------------------------
// One wave per 16 queries, both packed weight sets in LDS (96 KB per block: one block per CU).
// Forward (v_mfma_f32_16x16x32_bf16, rows = output neurons, columns = queries): wf fragment
// (L, t, b, lane) holds W_L[16t + (lane & 15)][k] with k = 32b + 8g + e for L = 0 and phi(...)
// for L > 0: the D fragment of a layer (neurons 16t + 4g + r of query j in lane j + 16g) is fed
// back as the next B operand. The BF16 activations a_L (after ReLU) stay in registers for the
// backward masks and are written transposed ([L][neuron][q], for the weight gradients).
// Loss: relative L2 (y - t)^2 / (y^2 + 0.01) summed over RGB and the batch (added to *loss),
// output gradient g_5 = bf16(2 (y - t) / (y^2 + 0.01) / n) for neurons 0..2.
// Data gradients: with z_L = W_L a_{L-1} (a_{-1} = x0), g_L = dLoss/dz_L follows
// g_{L-1}[q][i] = relu'(a_{L-1}[q][i]) * sum_o W_L[o][i] g_L[q][o]; wt fragment (L, t, b, lane)
// holds W_L[o][16t + (lane & 15)], o = phi(32b + 8g + e) for L < 5 and 32b + 8g + e for L = 5
// (g_5 enters in natural order, later g's as the previous D fragment). Each g_L is rounded to
// BF16 and written transposed (gT + L*64*ld, [o][q]; rows 3..63 of g_5 are never written and
// stay zero); layer 0 gives the input gradient dx = W_0^T g_0 in FP32 ([q][64]).
// ReLU derivative: active where the BF16 activation is > 0 (a -0 is inactive).
__global__ __launch_bounds__(256) void k_mlp_train(const uint4* __restrict__ wf, const uint4* __restrict__ wt,
                                                   const uint16_t* __restrict__ x0, const float* __restrict__ target,
                                                   const uint32_t* count, uint32_t nmax, uint32_t ld,
                                                   uint16_t* __restrict__ actsT, uint16_t* __restrict__ gT,
                                                   float* __restrict__ dx, float* __restrict__ loss) {
    const uint32_t n = batch_n(count, nmax);
    const uint32_t tiles = pad32(n) / 16u;
    if (blockIdx.x == 0 && threadIdx.x == 0) loss[kMaxLossBlocks] = __uint_as_float(n);  // the pass's samples
    if (blockIdx.x * 4u >= tiles) {  // no tile for this block's waves
        if (threadIdx.x == 0) loss[blockIdx.x] = 0.0f;
        return;
    }
    __shared__ uint4 s_wf[kLayers * 4 * 2 * 64];
    __shared__ float s_loss[4];
    __shared__ uint4 s_wt[kLayers * 4 * 2 * 64];
    fill_lds<kLayers * 4 * 2 * 64>(s_wf, wf);
    fill_lds<kLayers * 4 * 2 * 64>(s_wt, wt);
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, g = lane >> 4, j = lane & 15u;
    const uint32_t wave = (blockIdx.x * 256u + threadIdx.x) >> 6, nwaves = (gridDim.x * 256u) >> 6;
    const float inv_n = n ? 1.0f / (float)n : 0.0f;
    float lsum = 0.0f;
    for (uint32_t tile = wave; tile < tiles; tile += nwaves) {
        const uint32_t q = tile * 16u + j;
        const uint4* xr = reinterpret_cast<const uint4*>(x0 + (size_t)q * kW);
        bf16x8 B0 = as_frag(xr[g]), B1 = as_frag(xr[4 + g]);  // k = 8g.., 32 + 8g..
        uint32_t ha[kLayers - 1][4][2];  // a_L: BF16 pairs of neurons 16t + 4g + {0,1} and {2,3}
        f32x4 acc[4];
#pragma unroll
        for (int L = 0; L < kLayers; ++L) {
            const uint4* wl = s_wf + L * 512;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
                c = mfma(as_frag(wl[(t * 2 + 0) * 64 + lane]), B0, c);
                acc[t] = mfma(as_frag(wl[(t * 2 + 1) * 64 + lane]), B1, c);
            }
            if (L + 1 < kLayers) {
                uint16_t* aT = actsT + (size_t)L * kW * ld + q;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    ha[L][t][0] = pack2(fmaxf(acc[t][0], 0.0f), fmaxf(acc[t][1], 0.0f));
                    ha[L][t][1] = pack2(fmaxf(acc[t][2], 0.0f), fmaxf(acc[t][3], 0.0f));
                    const uint32_t nb = 16 * t + 4 * g;
                    aT[(size_t)(nb + 0) * ld] = (uint16_t)(ha[L][t][0] & 0xFFFFu);
                    aT[(size_t)(nb + 1) * ld] = (uint16_t)(ha[L][t][0] >> 16);
                    aT[(size_t)(nb + 2) * ld] = (uint16_t)(ha[L][t][1] & 0xFFFFu);
                    aT[(size_t)(nb + 3) * ld] = (uint16_t)(ha[L][t][1] >> 16);
                }
                B0 = as_frag(make_uint4(ha[L][0][0], ha[L][0][1], ha[L][1][0], ha[L][1][1]));
                B1 = as_frag(make_uint4(ha[L][2][0], ha[L][2][1], ha[L][3][0], ha[L][3][1]));
            }
        }
        // loss and output gradient: lanes g == 0 hold the outputs y[q][0..3] in acc[0]
        uint32_t v[3] = {0u, 0u, 0u};
        if (g == 0) {
            if (q < n) {
                for (int c = 0; c < 3; ++c) {
                    const float yc = acc[0][c], tc = target[(size_t)q * 4 + c];
                    const float den = yc * yc + 0.01f;
                    const float diff = yc - tc;
                    lsum += diff * diff / den;
                    v[c] = bf16_bits(((2.0f * diff) / den) * inv_n);
                }
            }
            uint16_t* g5T = gT + (size_t)(kLayers - 1) * kW * ld + q;
            for (int c = 0; c < 3; ++c) g5T[(size_t)c * ld] = (uint16_t)v[c];
        }
        B0 = as_frag(make_uint4(v[0] | (v[1] << 16), v[2], 0u, 0u));  // zero in lanes g > 0
        B1 = as_frag(make_uint4(0u, 0u, 0u, 0u));
#pragma unroll
        for (int L = kLayers - 1; L >= 0; --L) {
            const uint4* wl = s_wt + L * 512;
            f32x4 c[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                f32x4 z = {0.0f, 0.0f, 0.0f, 0.0f};
                z = mfma(as_frag(wl[(t * 2 + 0) * 64 + lane]), B0, z);
                c[t] = mfma(as_frag(wl[(t * 2 + 1) * 64 + lane]), B1, z);
            }
            if (L > 0) {
                uint16_t* out = gT + (size_t)(L - 1) * kW * ld + q;
                uint32_t h[4][2];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const uint32_t i0 = 16 * t + 4 * g;  // neurons i0..i0+3 of query q
                    const uint32_t m0 = ha[L - 1][t][0], m1 = ha[L - 1][t][1];
                    const float v0 = (m0 & 0x7FFFu) ? c[t][0] : 0.0f, v1 = ((m0 >> 16) & 0x7FFFu) ? c[t][1] : 0.0f;
                    const float v2 = (m1 & 0x7FFFu) ? c[t][2] : 0.0f, v3 = ((m1 >> 16) & 0x7FFFu) ? c[t][3] : 0.0f;
                    h[t][0] = pack2(v0, v1);
                    h[t][1] = pack2(v2, v3);
                    out[(size_t)(i0 + 0) * ld] = (uint16_t)(h[t][0] & 0xFFFFu);
                    out[(size_t)(i0 + 1) * ld] = (uint16_t)(h[t][0] >> 16);
                    out[(size_t)(i0 + 2) * ld] = (uint16_t)(h[t][1] & 0xFFFFu);
                    out[(size_t)(i0 + 3) * ld] = (uint16_t)(h[t][1] >> 16);
                }
                B0 = as_frag(make_uint4(h[0][0], h[0][1], h[1][0], h[1][1]));
                B1 = as_frag(make_uint4(h[2][0], h[2][1], h[3][0], h[3][1]));
            } else {
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    *reinterpret_cast<float4*>(dx + (size_t)q * kW + 16 * t + 4 * g) =
                        make_float4(c[t][0], c[t][1], c[t][2], c[t][3]);
            }
        }
    }
    // the block's loss sum, in a fixed order, to loss[block] (no atomics: one address shared by
    // every wave serializes them across the XCDs)
    for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
    if (lane == 0) s_loss[threadIdx.x >> 6] = lsum;
    __syncthreads();
    if (threadIdx.x == 0) loss[blockIdx.x] = ((s_loss[0] + s_loss[1]) + s_loss[2]) + s_loss[3];
}

// ---- weight gradients: dW_L[o][i] += sum_q g_L[q][o] a_{L-1}[q][i], layer L = blockIdx.y ----------
// A block takes kWgBlockRows queries of layer L; each of its four waves sums a quarter of them
// in MFMA accumulators (all 16 tiles of the 64x64 result), the four partial sums are added in
// LDS, and the block adds the result to dW_L with one atomic per weight.
struct WGradArgs {
    const uint16_t* gT[kLayers];  // g_L transposed, [o][q]
    const uint16_t* aT[kLayers];  // a_{L-1} transposed, [i][q] (x0 for L = 0)
};
constexpr uint32_t kWgWaveRows = 128, kWgBlockRows = 4 * kWgWaveRows;
__global__ __launch_bounds__(256) void k_bwd_weight(WGradArgs args, const uint32_t* count, uint32_t nmax, uint32_t ld,
                                                    float* __restrict__ gW) {
    const uint32_t npad = pad32(batch_n(count, nmax));
    if (blockIdx.x * kWgBlockRows >= npad) return;  // uniform over the block
    __shared__ float s_part[3][4096];
    const uint32_t L = blockIdx.y;
    const uint16_t* __restrict__ dT = args.gT[L];
    const uint16_t* __restrict__ aT = args.aT[L];
    float* __restrict__ dW = gW + (size_t)L * 4096u;
    const uint32_t lane = threadIdx.x & 63u, g = lane >> 4, j = lane & 15u, wv = threadIdx.x >> 6;
    const uint32_t q0 = blockIdx.x * kWgBlockRows + wv * kWgWaveRows;
    const uint32_t q1 = min(npad, q0 + kWgWaveRows);
    f32x4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t q = q0; q < q1; q += 32) {
        bf16x8 A[4], Bm[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            A[t] = as_frag(*reinterpret_cast<const uint4*>(dT + (size_t)(16 * t + j) * ld + q + 8 * g));
            Bm[t] = as_frag(*reinterpret_cast<const uint4*>(aT + (size_t)(16 * t + j) * ld + q + 8 * g));
        }
#pragma unroll
        for (int to = 0; to < 4; ++to)
#pragma unroll
            for (int ti = 0; ti < 4; ++ti) acc[to][ti] = mfma(A[to], Bm[ti], acc[to][ti]);
    }
    // element (to, ti, r) of lane (g, j) is dW[16to + 4g + r][16ti + j]
    if (wv > 0) {
#pragma unroll
        for (int to = 0; to < 4; ++to)
#pragma unroll
            for (int ti = 0; ti < 4; ++ti)
#pragma unroll
                for (int r = 0; r < 4; ++r) s_part[wv - 1][(16 * to + 4 * g + r) * kW + 16 * ti + j] = acc[to][ti][r];
    }
    __syncthreads();
    if (wv == 0) {
#pragma unroll
        for (int to = 0; to < 4; ++to)
#pragma unroll
            for (int ti = 0; ti < 4; ++ti)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const uint32_t k = (16 * to + 4 * g + r) * kW + 16 * ti + j;
                    atomicAdd(&dW[k], ((acc[to][ti][r] + s_part[0][k]) + s_part[1][k]) + s_part[2][k]);
                }
    }
}

// ---- hash-grid gradient by buckets --------------------------------------------------------------
// Sample q adds w_c * dL/dx0[2l + f] to feature f of each of its 8 corner entries c at level l
// (the adjoint of the trilinear encoding). Instead of a device-wide atomic per (sample, level,
// corner, feature), whose rate (~17 G/s) bound the training step, every contribution becomes a
// record in a bucket of 2^shift entries of one level, and one block
// per bucket (or per chunk of kGbChunk records of a large bucket) adds its records up in LDS and
// writes each touched entry once. Four launches per step (plus two scans):
//   count   per (256 samples, level) block, records per bucket (LDS histogram), bucket-major
//   scan    exclusive scans: per bucket over the blocks, then over the buckets (bucket and chunk
//           starts; a bucket of c records has ceil(c / kGbChunk) chunks)
//   scatter the records at their bucket offsets (rank inside the block from LDS counters)
//   accum   one block per chunk: LDS sums of the chunk's records; a bucket of one chunk is owned by
//           its block (plain read-add-write of each touched entry), a larger one flushes with
//           atomics (the coarse levels, whose few entries collect most records)
// Samples whose two gradient components of a level are zero leave no records.
constexpr uint32_t kGbChunk = 8192;       // records per accumulation block
constexpr uint32_t kGbMaxShift = 12;      // at most 4096 entries (32 KB of float pairs) per bucket

__device__ __forceinline__ bool grad_corners(const GridDev& g, const mrt_cache_sample* __restrict__ qs,
                                             const float* __restrict__ dx, uint32_t q, uint32_t l, uint32_t idx[8],
                                             float w[8], float2& gv) {
    gv = *reinterpret_cast<const float2*>(dx + (size_t)q * kW + 2 * l);
    if (gv.x == 0.0f && gv.y == 0.0f) return false;
    const float* qf = reinterpret_cast<const float*>(qs + q);
    const float pos[3] = {qf[0], qf[1], qf[2]};
    corners(g, l, pos, idx, w);
    return true;
}

__global__ __launch_bounds__(256) void k_gb_count(GridDev g, const mrt_cache_sample* __restrict__ qs,
                                                  const float* __restrict__ dx, const uint32_t* count, uint32_t nmax,
                                                  uint32_t shift, uint32_t bpl, uint32_t* __restrict__ hist) {
    extern __shared__ uint32_t s_h[];  // bpl counters
    const uint32_t l = blockIdx.y, bx = blockIdx.x, nbx = gridDim.x, tid = threadIdx.x;
    for (uint32_t i = tid; i < bpl; i += 256u) s_h[i] = 0u;
    __syncthreads();
    const uint32_t q = bx * 256u + tid;
    uint32_t idx[8];
    float w[8];
    float2 gv;
    if (q < batch_n(count, nmax) && grad_corners(g, qs, dx, q, l, idx, w, gv)) {
#pragma unroll
        for (int c = 0; c < 8; ++c) atomicAdd(&s_h[idx[c] >> shift], 1u);
    }
    __syncthreads();
    for (uint32_t i = tid; i < bpl; i += 256u) hist[(size_t)(l * bpl + i) * nbx + bx] = s_h[i];
}

// In-place exclusive scan of each bucket's row of block counts (one block per bucket).
__global__ __launch_bounds__(256) void k_gb_scan_rows(uint32_t* __restrict__ hist, uint32_t nbx,
                                                      uint32_t* __restrict__ row_total) {
    __shared__ uint32_t s[256];
    __shared__ uint32_t s_total;
    uint32_t* row = hist + (size_t)blockIdx.x * nbx;
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nbx; b0 += 256u) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t v = i < nbx ? row[i] : 0u;
        s[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t off = 1; off < 256u; off <<= 1) {
            const uint32_t add = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
            __syncthreads();
            s[threadIdx.x] += add;
            __syncthreads();
        }
        const uint32_t incl = s[threadIdx.x];
        if (i < nbx) row[i] = carry + incl - v;
        if (threadIdx.x == 255u) s_total = incl;
        __syncthreads();
        carry += s_total;
        __syncthreads();
    }
    if (threadIdx.x == 0) row_total[blockIdx.x] = carry;
}

// Exclusive scans over the buckets (one block of 1024 threads): bucket starts and chunk starts,
// with the totals at [nb], and every chunk's descriptor {bucket, first record, end, owner}.
__global__ __launch_bounds__(1024) void k_gb_scan_buckets(const uint32_t* __restrict__ row_total, uint32_t nb,
                                                          uint32_t* __restrict__ bstart, uint32_t* __restrict__ cstart,
                                                          uint4* __restrict__ cdesc) {
    __shared__ uint32_t s_r[1024], s_c[1024];
    const uint32_t tid = threadIdx.x;
    const uint32_t per = (nb + 1023u) / 1024u, b0 = tid * per;
    uint32_t r = 0, c = 0;
    for (uint32_t i = 0; i < per; ++i) {
        const uint32_t b = b0 + i;
        if (b < nb) { const uint32_t v = row_total[b]; r += v; c += (v + kGbChunk - 1u) / kGbChunk; }
    }
    s_r[tid] = r;
    s_c[tid] = c;
    __syncthreads();
    for (uint32_t off = 1; off < 1024u; off <<= 1) {
        const uint32_t ar = tid >= off ? s_r[tid - off] : 0u, ac = tid >= off ? s_c[tid - off] : 0u;
        __syncthreads();
        s_r[tid] += ar;
        s_c[tid] += ac;
        __syncthreads();
    }
    uint32_t rr = s_r[tid] - r, cc = s_c[tid] - c;  // exclusive
    for (uint32_t i = 0; i < per; ++i) {
        const uint32_t b = b0 + i;
        if (b < nb) {
            bstart[b] = rr;
            cstart[b] = cc;
            const uint32_t v = row_total[b], nc = (v + kGbChunk - 1u) / kGbChunk;
            for (uint32_t k = 0; k < nc; ++k)
                cdesc[cc + k] = make_uint4(b, rr + k * kGbChunk, min(rr + v, rr + (k + 1u) * kGbChunk), nc == 1u ? 1u : 0u);
            rr += v;
            cc += nc;
        }
    }
    if (tid == 1023u) { bstart[nb] = s_r[1023]; cstart[nb] = s_c[1023]; }
}

__global__ __launch_bounds__(256) void k_gb_scatter(GridDev g, const mrt_cache_sample* __restrict__ qs,
                                                    const float* __restrict__ dx, const uint32_t* count, uint32_t nmax,
                                                    uint32_t shift, uint32_t bpl, const uint32_t* __restrict__ hist,
                                                    const uint32_t* __restrict__ bstart, uint16_t* __restrict__ rec_idx,
                                                    float2* __restrict__ rec_val) {
    extern __shared__ uint32_t s_n[];  // bpl next positions: the bucket's start + this block's offset in it
    const uint32_t l = blockIdx.y, bx = blockIdx.x, nbx = gridDim.x, tid = threadIdx.x;
    for (uint32_t i = tid; i < bpl; i += 256u) s_n[i] = bstart[l * bpl + i] + hist[(size_t)(l * bpl + i) * nbx + bx];
    __syncthreads();
    const uint32_t q = bx * 256u + tid;
    uint32_t idx[8];
    float w[8];
    float2 gv;
    if (q < batch_n(count, nmax) && grad_corners(g, qs, dx, q, l, idx, w, gv)) {
        const uint32_t mask = (1u << shift) - 1u;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const uint32_t pos = atomicAdd(&s_n[idx[c] >> shift], 1u);
            rec_idx[pos] = (uint16_t)(idx[c] & mask);
            rec_val[pos] = make_float2(w[c] * gv.x, w[c] * gv.y);
        }
    }
}

// clean != 0: the table gradient is zero on entry (an Adam step consumed the last one), so an
// owning block stores 0 + sum without reading the entry. Otherwise it reads the bucket's entries
// first, all loads in flight together, then adds and stores.
__global__ __launch_bounds__(256) void k_gb_accum(const uint16_t* __restrict__ rec_idx, const float2* __restrict__ rec_val,
                                                  const uint32_t* __restrict__ cstart, const uint4* __restrict__ cdesc,
                                                  uint32_t nb, uint32_t shift, uint32_t bpl, uint32_t log2_table,
                                                  uint32_t clean, float* __restrict__ gtable) {
    __shared__ __align__(16) float s_acc[2u << kGbMaxShift];
    const uint32_t k = blockIdx.x, tid = threadIdx.x;
    const uint32_t nchunks = cstart[nb];
    const uint4 d = cdesc[k];  // in the allocation for every launched block; valid for k < nchunks
    if (k >= nchunks) return;
    const uint32_t b = d.x, r0 = d.y, r1 = d.z;
    const bool owner = d.w != 0u;
    const uint32_t ent = 1u << shift;
    for (uint32_t i = tid; i < ent / 2u; i += 256u) reinterpret_cast<float4*>(s_acc)[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    __syncthreads();
    // The LDS float atomics are the kernel's main cost (about half of it at the workload's batch);
    // larger chunks lengthen the coarse levels' blocks, which end the kernel.
    for (uint32_t r = r0 + tid; r < r1; r += 1024u) {  // four records per thread in flight
        uint32_t e[4];
        float2 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t rr = r + 256u * u;
            e[u] = rr < r1 ? (uint32_t)rec_idx[rr] : 0u;
            v[u] = rr < r1 ? rec_val[rr] : make_float2(0.0f, 0.0f);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (r + 256u * u < r1) {
                atomicAdd(&s_acc[2u * e[u]], v[u].x);
                atomicAdd(&s_acc[2u * e[u] + 1u], v[u].y);
            }
    }
    __syncthreads();
    const uint32_t l = b / bpl, bl = b % bpl;
    float2* base = reinterpret_cast<float2*>(gtable + (((size_t)l << log2_table) + ((size_t)bl << shift)) * 2u);
    constexpr uint32_t kPer = (1u << kGbMaxShift) / 256u;  // entries per thread at the largest bucket
    if (owner && clean) {  // stores only
#pragma unroll
        for (uint32_t i = 0; i < kPer; ++i) {
            const uint32_t e = tid + 256u * i;
            if (e >= ent) continue;
            const float ax = s_acc[2u * e], ay = s_acc[2u * e + 1u];
            if (ax != 0.0f || ay != 0.0f) base[e] = make_float2(0.0f + ax, 0.0f + ay);
        }
    } else if (owner) {
        float2 old[kPer];
#pragma unroll
        for (uint32_t i = 0; i < kPer; ++i)
            old[i] = tid + 256u * i < ent ? base[tid + 256u * i] : make_float2(0.0f, 0.0f);
#pragma unroll
        for (uint32_t i = 0; i < kPer; ++i) {
            const uint32_t e = tid + 256u * i;
            if (e >= ent) continue;
            const float ax = s_acc[2u * e], ay = s_acc[2u * e + 1u];
            if (ax != 0.0f || ay != 0.0f) base[e] = make_float2(old[i].x + ax, old[i].y + ay);
        }
    } else {
        for (uint32_t e = tid; e < ent; e += 256u) {
            const float ax = s_acc[2u * e], ay = s_acc[2u * e + 1u];
            if (ax == 0.0f && ay == 0.0f) continue;
            atomicAdd(&base[e].x, ax);
            atomicAdd(&base[e].y, ay);
        }
    }
}

// ---- Adam (Kingma & Ba 2015) with bias correction; clears the gradient ------------------------
// sparse != 0 (hash table): an entry whose gradient is exactly zero is not updated at all, its
// moments included (Mueller et al. 2022), so an entry that one sample touched moves by one step
// and does not drift on its decaying momentum. Dense (MLP weights): the textbook update; an
// entry whose gradient and moments are all zero is left as that update would leave it.
// With ema != null, ema = a * ema + (1 - a) * p after the update (exponential average of the
// parameters, used for inference with bias correction 1 / (1 - a^t)).
// The table's average over a group of S consecutive steps (a frame's training steps) is applied
// in closed form: the group's first step moves every entry to a^S ema + (1 - a^S) p, and step j
// (1-based) adds (1 - a^(S-j+1)) (p_new - p_old) to each entry it updates, which equals S
// per-step averages in exact arithmetic; the S - 1 later steps then read the gradient alone of
// the entries they leave. A group of one step uses the per-step formula.
struct AdamHyper {
    float b1, b2, eps, c1, c2, ema_a;
    float ema_aS, ema_gain;  // a^S and 1 - a^(S-j+1) of the table's group
    uint32_t ema_base;       // the group's first step
    uint32_t ema_exact;      // a group of one step
};

// One element: returns the new parameter (p itself when skipped).
__device__ __forceinline__ float adam_elem(float p, float gi, float& m, float& v, bool& updated, float lr,
                                           const AdamHyper& h, int sparse) {
    updated = !(sparse ? gi == 0.0f : (gi == 0.0f && m == 0.0f && v == 0.0f));
    if (!updated) return p;
    const float mi = h.b1 * m + (1.0f - h.b1) * gi;
    const float vi = h.b2 * v + (1.0f - h.b2) * (gi * gi);
    m = mi;
    v = vi;
    const float mh = mi / h.c1, vh = vi / h.c2;
    return p - lr * mh / (sqrtf(vh) + h.eps);
}

// Dense or sparse Adam over [begin, n) with the given stride (scalar accesses).
__device__ void adam_range(float* __restrict__ p, float* __restrict__ grad, float* __restrict__ m,
                           float* __restrict__ v, float* __restrict__ ema, size_t n, size_t begin, size_t stride,
                           float lr, const AdamHyper& h, int sparse) {
    for (size_t i = begin; i < n; i += stride) {
        float mi = m[i], vi = v[i];
        bool upd;
        const float pn = adam_elem(p[i], grad[i], mi, vi, upd, lr, h, sparse);
        if (upd) {
            m[i] = mi;
            v[i] = vi;
            p[i] = pn;
            grad[i] = 0.0f;
        }
        if (ema) ema[i] = h.ema_a * ema[i] + (1.0f - h.ema_a) * pn;
    }
}

// ---- weight packing ------------------------------------------------------------------------
// wf (forward A fragments): layer L, tile t, block b, lane l, element e -> W_L[16t + (l&15)][k]
// with k = 32b + 8(l>>4) + e for L = 0 and phi(32b + 8(l>>4) + e) for L > 0.
// wt (data-gradient A fragments): -> W_L[o][16t + (l&15)] with o = phi(32b + 8(l>>4) + e) for
// L < 5 and o = 32b + 8(l>>4) + e for L = 5.
// wq (inference A fragments, 32x32x16): layer L, tile T, step s, lane l, element e ->
// W_L[32T + (l&31)][k] with k = 16s + 8(l>>5) + e for L = 0 and phi32(16s + 8(l>>5) + e) for L > 0.
// Each output is optional (null: not written); i indexes the L * 4096 packed elements.
__device__ __forceinline__ void pack_elem(uint32_t i, const float* __restrict__ W, float scale, uint16_t* __restrict__ wf,
                                          uint16_t* __restrict__ wt, uint16_t* __restrict__ wq) {
    const uint32_t L = i >> 12, r = i & 4095u;
    const float* w = W + (size_t)L * 4096u;
    {
        const uint32_t e = r & 7u, lane = (r >> 3) & 63u, b = (r >> 9) & 1u, t = r >> 10;
        const uint32_t row = 16 * t + (lane & 15u);
        const int slot = (int)(32 * b + 8 * (lane >> 4) + e);
        const int k = L == 0 ? slot : phi(slot);
        if (wf) wf[i] = (uint16_t)bf16_bits(w[row * kW + k] * scale);
        const int ko = L + 1 < (uint32_t)kLayers ? phi(slot) : slot;  // K order of k_mlp_train's chain
        if (wt) wt[i] = (uint16_t)bf16_bits(w[ko * kW + row] * scale);
    }
    if (wq) {
        const uint32_t e = r & 7u, lane = (r >> 3) & 63u, s = (r >> 9) & 3u, T = r >> 11;
        const uint32_t row = 32 * T + (lane & 31u);
        const int slot = (int)(16 * s + 8 * (lane >> 5) + e);
        const int k = L == 0 ? slot : phi32(slot);
        wq[i] = (uint16_t)bf16_bits(w[row * kW + k] * scale);
    }
}

// wf and wt from W; wq from Wq * q_scale (the weights, or the bias-corrected running average).
__global__ __launch_bounds__(256) void k_pack(const float* __restrict__ W, uint16_t* __restrict__ wf,
                                              uint16_t* __restrict__ wt, const float* __restrict__ Wq, float q_scale,
                                              uint16_t* __restrict__ wq) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;  // over L * 4096 elements
    if (i >= (uint32_t)kLayers * 4096u) return;
    pack_elem(i, W, 1.0f, wf, wt, Wq == W && q_scale == 1.0f ? wq : nullptr);
    if (Wq != W || q_scale != 1.0f) pack_elem(i, Wq, q_scale, nullptr, nullptr, wq);
}

// ---- one optimizer step: the MLP weights in blocks 0..kWeightBlocks-1, the table in the rest ----
// Dense Adam for the weights (one element per thread); sparse Adam over the table four entries
// at a time: a group of four untouched entries reads neither moments nor parameters beyond the
// group's gradient, except on a group's first step, where the running average (if any) moves
// and every parameter is checked.
constexpr uint32_t kWeightBlocks = kLayers * 4096 / 256;
// Table groups of four floats the encoding can read: [begin[r], end[r]) for r < n. A dense
// level's slots past its first (res+1)^3 entries (rounded up to an even count) are left out.
struct TableRanges {
    uint32_t n;
    uint32_t begin[kMaxLevels], end[kMaxLevels];
};
struct AdamStep {
    float *W, *gW, *mW, *vW, *eW;               // weights (eW: average or null)
    float *T, *gT, *mT, *vT, *eT;               // table (eT: average or null); table_floats % 4 == 0
    size_t table_floats;
    float lr_w, lr_t;
    uint32_t* flag;       // |= 1 when a parameter is not finite after the step
    uint32_t* flag_next;  // the previous step's flag; cleared for the next step
    uint32_t carry;       // not a group's first step: the previous verdict is carried over (the
                          // entries this step leaves are as the previous step checked them)
    TableRanges used;     // the table groups to update
};
__global__ __launch_bounds__(256) void k_adam_step(AdamStep a, AdamHyper h) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const uint32_t prev = *a.flag_next;
        *a.flag_next = 0u;
        if (a.carry && prev) atomicOr(a.flag, 1u);
    }
    bool bad = false;
    if (blockIdx.x < kWeightBlocks) {
        const size_t i = blockIdx.x * 256u + threadIdx.x;
        adam_range(a.W, a.gW, a.mW, a.vW, a.eW, (size_t)kLayers * 4096, i, (size_t)kLayers * 4096, a.lr_w, h, 0);
        bad = !isfinite(a.W[i]) || (a.eW && !isfinite(a.eW[i]));  // queries read the average when it is on
        if (__ballot(bad) != 0ull && __lane_id() == 0) atomicOr(a.flag, 1u);
        return;
    }
    const size_t stride = (size_t)(gridDim.x - kWeightBlocks) * 256u;
    float4* __restrict__ p4 = reinterpret_cast<float4*>(a.T);
    float4* __restrict__ g4 = reinterpret_cast<float4*>(a.gT);
    float4* __restrict__ m4 = reinterpret_cast<float4*>(a.mT);
    float4* __restrict__ v4 = reinterpret_cast<float4*>(a.vT);
    float4* __restrict__ e4 = reinterpret_cast<float4*>(a.eT);
    for (uint32_t r = 0; r < a.used.n; ++r)  // wave-uniform: range bounds are kernel arguments
    for (size_t i = a.used.begin[r] + (size_t)(blockIdx.x - kWeightBlocks) * 256u + threadIdx.x; i < a.used.end[r];
         i += stride) {
        const float4 g = g4[i];
        const bool untouched = g.x == 0.0f && g.y == 0.0f && g.z == 0.0f && g.w == 0.0f;
        if (untouched && !h.ema_base) continue;
        const float4 p = p4[i];
        float4 e = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (h.ema_base) {
            bad |= !(isfinite(p.x) && isfinite(p.y) && isfinite(p.z) && isfinite(p.w));
            if (e4) {
                e = e4[i];
                const float b = h.ema_exact ? h.ema_a : h.ema_aS;
                if (!(h.ema_exact && !untouched)) {  // a touched entry of a one-step group: below
                    e.x = b * e.x + (1.0f - b) * p.x;
                    e.y = b * e.y + (1.0f - b) * p.y;
                    e.z = b * e.z + (1.0f - b) * p.z;
                    e.w = b * e.w + (1.0f - b) * p.w;
                }
                if (untouched) {
                    e4[i] = e;
                    bad |= !(isfinite(e.x) && isfinite(e.y) && isfinite(e.z) && isfinite(e.w));
                }
            }
            if (untouched) continue;
        } else if (e4) {
            e = e4[i];
        }
        float4 m = m4[i], v = v4[i], pn;
        bool u0, u1, u2, u3;
        pn.x = adam_elem(p.x, g.x, m.x, v.x, u0, a.lr_t, h, 1);
        pn.y = adam_elem(p.y, g.y, m.y, v.y, u1, a.lr_t, h, 1);
        pn.z = adam_elem(p.z, g.z, m.z, v.z, u2, a.lr_t, h, 1);
        pn.w = adam_elem(p.w, g.w, m.w, v.w, u3, a.lr_t, h, 1);
        m4[i] = m;  // components left alone are written back as read
        v4[i] = v;
        p4[i] = pn;
        bad |= !(isfinite(pn.x) && isfinite(pn.y) && isfinite(pn.z) && isfinite(pn.w));
        g4[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (e4) {
            if (h.ema_exact) {
                e.x = h.ema_a * e.x + (1.0f - h.ema_a) * pn.x;
                e.y = h.ema_a * e.y + (1.0f - h.ema_a) * pn.y;
                e.z = h.ema_a * e.z + (1.0f - h.ema_a) * pn.z;
                e.w = h.ema_a * e.w + (1.0f - h.ema_a) * pn.w;
            } else {  // a component left alone has pn == p and adds zero
                e.x += h.ema_gain * (pn.x - p.x);
                e.y += h.ema_gain * (pn.y - p.y);
                e.z += h.ema_gain * (pn.z - p.z);
                e.w += h.ema_gain * (pn.w - p.w);
            }
            e4[i] = e;
            bad |= !(isfinite(e.x) && isfinite(e.y) && isfinite(e.z) && isfinite(e.w));
        }
    }
    if (__ballot(bad) != 0ull && __lane_id() == 0) atomicOr(a.flag, 1u);
}

__global__ __launch_bounds__(256) void k_bf16_to_f32(const uint16_t* __restrict__ in, float* __restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i < n) out[i] = bf16_float(in[i]);
}

__global__ __launch_bounds__(256) void k_zero_f32(float* p, size_t n) {
    const size_t stride = (size_t)gridDim.x * 256u;
    for (size_t i = (size_t)blockIdx.x * 256u + threadIdx.x; i < n; i += stride) p[i] = 0.0f;
}

// dst = src * s (restarting a running average at given parameters); s = 0 writes +0 (an average
// before the first step is zero, whatever the sign of the parameter).
__global__ __launch_bounds__(256) void k_scale_copy(float* __restrict__ dst, const float* __restrict__ src, size_t n, float s) {
    const size_t stride = (size_t)gridDim.x * 256u;
    for (size_t i = (size_t)blockIdx.x * 256u + threadIdx.x; i < n; i += stride) dst[i] = s == 0.0f ? 0.0f : src[i] * s;
}

// flag[0] |= 1 if any of the n values is not finite.
__global__ __launch_bounds__(256) void k_check_finite(const float* __restrict__ v, size_t n, uint32_t* flag) {
    const size_t stride = (size_t)gridDim.x * 256u;
    bool bad = false;
    for (size_t i = (size_t)blockIdx.x * 256u + threadIdx.x; i < n; i += stride) bad |= !isfinite(v[i]);
    if (__ballot(bad) != 0ull && __lane_id() == 0) atomicOr(flag, 1u);
}

// flag[0] |= 1 if a table entry in the used ranges (groups of four floats) is not finite.
__global__ __launch_bounds__(256) void k_check_finite_ranges(const float4* __restrict__ t, TableRanges used,
                                                             uint32_t* flag) {
    const size_t stride = (size_t)gridDim.x * 256u;
    bool bad = false;
    for (uint32_t r = 0; r < used.n; ++r)
        for (size_t i = used.begin[r] + (size_t)blockIdx.x * 256u + threadIdx.x; i < used.end[r]; i += stride) {
            const float4 p = t[i];
            bad |= !(isfinite(p.x) && isfinite(p.y) && isfinite(p.z) && isfinite(p.w));
        }
    if (__ballot(bad) != 0ull && __lane_id() == 0) atomicOr(flag, 1u);
}

inline uint32_t blocks_for(size_t n) { return (uint32_t)((n + 255) / 256); }

}  // namespace

// ---- the cache object ---------------------------------------------------------------------------
struct CacheScratch {
    size_t cap = 0;  // rows (multiple of 32); also the leading dimension of the transposed copies
    uint16_t *x0 = nullptr, *x0T = nullptr, *actsT = nullptr;
    uint16_t *gT = nullptr;  // g_0..g_5 transposed, [L][64][ld] (rows 3..63 of g_5 stay zero)
    float *y = nullptr, *dx = nullptr;
    // table-gradient buckets (training): records, block counts [bucket][block], bucket totals,
    // bucket and chunk starts
    uint16_t* rec_idx = nullptr;
    float2* rec_val = nullptr;
    uint32_t *gb_hist = nullptr, *gb_rows = nullptr, *gb_bstart = nullptr, *gb_cstart = nullptr;
    uint4* gb_cdesc = nullptr;  // descriptor of each accumulation chunk
};

struct mrt_cache {
    GridDev grid{};
    mrt_cache_opts opts{};
    size_t table_floats = 0;
    float* d_W = nullptr;       // 6*64*64 master weights
    float* d_gW = nullptr;
    float* d_mW = nullptr;
    float* d_vW = nullptr;
    float* d_table = nullptr;   // levels * 2^T * 2
    float* d_gtable = nullptr;
    float* d_mtable = nullptr;
    float* d_vtable = nullptr;
    uint16_t* d_wf = nullptr;   // packed forward fragments
    uint16_t* d_wt = nullptr;   // packed data-gradient fragments
    float* d_eW = nullptr;      // exponential averages of the weights and the table (ema_decay > 0)
    float* d_etable = nullptr;
    uint16_t* d_wq = nullptr;   // packed inference fragments (32x32x16 layout) of the weights queries
                                // use: the bias-corrected average when averaging is on, else d_W
    float* d_loss = nullptr;    // [b] loss sum of block b of the last training pass's k_mlp_train,
                                // [kMaxLossBlocks] its sample count (uint32 bits)
    uint32_t loss_blocks = 0;   // its blocks
    uint32_t* d_flag = nullptr; // [0]: 1 if a weight or a used table entry is not finite (set by a full check);
                                // [1 + (s & 1)]: the same, found by Adam step s while updating
    uint32_t step = 0;
    bool step_flag_current = false;  // the last Adam step read every parameter and nothing changed them since
    bool grad_clean = true;          // the table gradient is all zero (creation, or an Adam step since the last pass)
    CacheScratch inf;           // inference batches (encoding + outputs)
    CacheScratch trn;           // training batches (activations and deltas kept for the backward)
    uint64_t device_bytes = 0;
};

static void scratch_free(CacheScratch& s) {
    hipFree(s.x0); hipFree(s.x0T); hipFree(s.actsT); hipFree(s.gT);
    hipFree(s.y); hipFree(s.dx);
    hipFree(s.rec_idx); hipFree(s.rec_val); hipFree(s.gb_hist); hipFree(s.gb_rows); hipFree(s.gb_bstart);
    hipFree(s.gb_cstart); hipFree(s.gb_cdesc);
    s = CacheScratch{};
}

static void cache_free(mrt_cache* c) {
    if (!c) return;
    scratch_free(c->inf);
    scratch_free(c->trn);
    hipFree(c->d_W); hipFree(c->d_gW); hipFree(c->d_mW); hipFree(c->d_vW);
    hipFree(c->d_table); hipFree(c->d_gtable); hipFree(c->d_mtable); hipFree(c->d_vtable);
    hipFree(c->d_wf); hipFree(c->d_wt); hipFree(c->d_loss); hipFree(c->d_flag);
    hipFree(c->d_eW); hipFree(c->d_etable); hipFree(c->d_wq);
    delete c;
}

// Buckets of the table gradient: 2^shift entries each, bpl per level, nb in all.
struct GbShape { uint32_t shift, bpl, nb; };
static GbShape gb_shape(const GridDev& g) {
    GbShape b;
    b.shift = std::min<uint32_t>(kGbMaxShift, g.log2_table);
    b.bpl = 1u << (g.log2_table - b.shift);
    b.nb = g.levels * b.bpl;
    return b;
}

static hipError_t scratch_ensure(CacheScratch& s, size_t n, bool training, const GridDev& g) {
    const size_t npad = (std::max<size_t>(n, 1) + 31) & ~(size_t)31;
    if (npad <= s.cap) return hipSuccess;
    scratch_free(s);
    hipError_t e = hipSuccess;
    auto A = [&](auto** p, size_t bytes) { if (e == hipSuccess) e = hipMalloc((void**)p, bytes); };
    A(&s.x0, npad * kW * 2);
    A(&s.y, npad * 16);
    if (training) {
        A(&s.x0T, npad * kW * 2);
        A(&s.actsT, (size_t)(kLayers - 1) * npad * kW * 2);
        A(&s.gT, (size_t)kLayers * npad * kW * 2);
        A(&s.dx, npad * kW * 4);
        const GbShape b = gb_shape(g);
        const size_t recs = npad * g.levels * 8, nbx = (npad + 255) / 256;
        A(&s.rec_idx, recs * 2);
        A(&s.rec_val, recs * 8);
        A(&s.gb_hist, (size_t)b.nb * nbx * 4);
        A(&s.gb_rows, (size_t)b.nb * 4);
        A(&s.gb_bstart, ((size_t)b.nb + 1) * 4);
        A(&s.gb_cstart, ((size_t)b.nb + 1) * 4);
        A(&s.gb_cdesc, ((size_t)b.nb + (recs + kGbChunk - 1) / kGbChunk) * 16);
        if (e == hipSuccess) e = hipMemset(s.gT, 0, (size_t)kLayers * npad * kW * 2);
        if (e == hipSuccess) e = hipDeviceSynchronize();
    }
    if (e != hipSuccess) { scratch_free(s); return e; }
    s.cap = npad;
    return hipSuccess;
}

static bool ema_on(const mrt_cache* c) { return c->opts.ema_decay > 0.0f && c->step > 0; }
static float ema_scale(const mrt_cache* c) {
    return (float)(1.0 / (1.0 - std::pow((double)c->opts.ema_decay, (double)c->step)));
}

static hipError_t cache_pack(mrt_ctx* ctx, mrt_cache* c) {
    const bool avg = ema_on(c);
    hipLaunchKernelGGL(k_pack, dim3(blocks_for((size_t)kLayers * 4096)), dim3(256), 0, ctx->stream, c->d_W, c->d_wf,
                       c->d_wt, avg ? c->d_eW : c->d_W, avg ? ema_scale(c) : 1.0f, c->d_wq);
    return hipGetLastError();
}

// k_mlp_infer: 64 queries per wave and iteration, 4 waves per block, 48 KB of LDS per block
// (three blocks fit one CU's 160 KB).
static uint32_t infer_blocks(mrt_ctx* ctx, size_t npad) {
    const size_t iters = (npad + 63) / 64;
    return (uint32_t)std::max<size_t>(1, std::min<size_t>((iters + 3) / 4, (size_t)ctx->num_cus * 3));
}

static size_t pad32h(size_t n) { return (n + 31) & ~(size_t)31; }

// Encoding of min(*count, nmax) samples into s.x0 (and x0T when given): one thread per (query,
// level) while one thread per query would fill fewer than 1024 threads per CU.
// clear (optional): a float set to zero by the launch (the training step's loss sum).
static void launch_encode(mrt_ctx* ctx, const mrt_cache* c, const float* table, float scale, const mrt_cache_sample* qs,
                          const uint32_t* count, uint32_t nmax, uint32_t ld, uint16_t* x0, uint16_t* x0T,
                          float* clear = nullptr) {
    const size_t rows = pad32h(nmax);
    if (rows < (size_t)ctx->num_cus * 1024)
        hipLaunchKernelGGL(k_encode_lv, dim3(blocks_for(rows * kMaxLevels)), dim3(256), 0, ctx->stream, c->grid, table,
                           scale, qs, count, nmax, ld, x0, x0T, clear);
    else
        hipLaunchKernelGGL(k_encode, dim3(blocks_for(rows)), dim3(256), 0, ctx->stream, c->grid, table, scale, qs, count,
                           nmax, ld, x0, x0T, clear);
}

// Queries: encoding + inference forward over min(*count, nmax) samples (count may be null: nmax
// samples), into s.y.
static hipError_t cache_forward(mrt_ctx* ctx, mrt_cache* c, CacheScratch& s, const mrt_cache_sample* qs,
                                const uint32_t* count, uint32_t nmax) {
    const bool avg = ema_on(c);  // queries use the bias-corrected averages
    {
        ProfScope p(ctx, "cache_encode");
        launch_encode(ctx, c, avg ? c->d_etable : c->d_table, avg ? ema_scale(c) : 1.0f, qs, count, nmax, (uint32_t)s.cap,
                      s.x0, nullptr);
    }
    {
        ProfScope p(ctx, "cache_mlp_forward");
        hipLaunchKernelGGL(k_mlp_infer, dim3(infer_blocks(ctx, pad32h(nmax))), dim3(256), 0, ctx->stream,
                           reinterpret_cast<const uint4*>(c->d_wq), s.x0, count, nmax, s.y);
    }
    return hipGetLastError();
}

// One training pass over min(*count, nmax) samples: encoding, forward + loss + data gradients,
// weight gradients, hash-grid gradient; gradients accumulate into d_gW / d_gtable.
static hipError_t cache_train_pass(mrt_ctx* ctx, mrt_cache* c, const mrt_cache_sample* qs, const float* target,
                                   const uint32_t* count, uint32_t nmax) {
    CacheScratch& s = c->trn;
    const uint32_t ld = (uint32_t)s.cap;
    const size_t npad = pad32h(nmax);
    {
        ProfScope p(ctx, "cache_encode");
        launch_encode(ctx, c, c->d_table, 1.0f, qs, count, nmax, ld, s.x0, s.x0T);
    }
    {
        ProfScope p(ctx, "cache_mlp_train");  // the MLP's forward, data and weight gradients
        const uint32_t tblocks = (uint32_t)std::max<size_t>(
            1, std::min<size_t>({(npad / 16 + 3) / 4, (size_t)ctx->num_cus, (size_t)kMaxLossBlocks}));
        hipLaunchKernelGGL(k_mlp_train, dim3(tblocks), dim3(256), 0, ctx->stream, reinterpret_cast<const uint4*>(c->d_wf),
                           reinterpret_cast<const uint4*>(c->d_wt), s.x0, target, count, nmax, ld, s.actsT, s.gT, s.dx,
                           c->d_loss);
        c->loss_blocks = tblocks;
        WGradArgs wa;
        for (int L = 0; L < kLayers; ++L) {
            wa.gT[L] = s.gT + (size_t)L * kW * ld;
            wa.aT[L] = L > 0 ? s.actsT + (size_t)(L - 1) * ld * kW : s.x0T;
        }
        const uint32_t wblocks = (uint32_t)((npad + kWgBlockRows - 1) / kWgBlockRows);
        hipLaunchKernelGGL(k_bwd_weight, dim3(wblocks, kLayers), dim3(256), 0, ctx->stream, wa, count, nmax, ld, c->d_gW);
    }
    {
        ProfScope pg(ctx, "cache_grid_grad");
        const GbShape b = gb_shape(c->grid);
        const uint32_t nbx = (uint32_t)((npad + 255) / 256);
        const dim3 grid(nbx, c->grid.levels);
        hipLaunchKernelGGL(k_gb_count, grid, dim3(256), b.bpl * 4, ctx->stream, c->grid, qs, s.dx, count, nmax, b.shift,
                           b.bpl, s.gb_hist);
        hipLaunchKernelGGL(k_gb_scan_rows, dim3(b.nb), dim3(256), 0, ctx->stream, s.gb_hist, nbx, s.gb_rows);
        hipLaunchKernelGGL(k_gb_scan_buckets, dim3(1), dim3(1024), 0, ctx->stream, s.gb_rows, b.nb, s.gb_bstart,
                           s.gb_cstart, s.gb_cdesc);
        hipLaunchKernelGGL(k_gb_scatter, grid, dim3(256), b.bpl * 4, ctx->stream, c->grid, qs, s.dx, count, nmax, b.shift,
                           b.bpl, s.gb_hist, s.gb_bstart, s.rec_idx, s.rec_val);
        // chunks: at most one partial chunk per bucket plus the full ones
        const size_t chunks = (size_t)b.nb + (npad * c->grid.levels * 8 + kGbChunk - 1) / kGbChunk;
        hipLaunchKernelGGL(k_gb_accum, dim3((uint32_t)chunks), dim3(256), 0, ctx->stream, s.rec_idx, s.rec_val,
                           s.gb_cstart, s.gb_cdesc, b.nb, b.shift, b.bpl, c->grid.log2_table,
                           c->grad_clean ? 1u : 0u, c->d_gtable);
    }
    c->grad_clean = false;
    return hipGetLastError();
}

// pos, len: the step's place in a group of len steps whose table average is applied in closed form.

// The table's used groups: a dense level's (res+1)^3 entries, a hashed level's 2^T; adjacent
// ranges merged. These are the entries the encoding can read; the Adam step updates and checks
// them, and so does the full non-finite check.
static TableRanges table_ranges(const mrt_cache* c) {
    TableRanges u;
    const uint64_t tsize4 = (1ull << c->grid.log2_table) / 2;  // groups of 4 floats = 2 entries
    u.n = 0;
    for (uint32_t l = 0; l < c->grid.levels; ++l) {
        const uint64_t side = (uint64_t)c->grid.res[l] + 1;
        const uint64_t used = c->grid.dense[l] ? (side * side * side + 1) / 2 : tsize4;
        const uint64_t b0 = l * tsize4, b1 = b0 + used;
        if (u.n && u.end[u.n - 1] == b0) u.end[u.n - 1] = (uint32_t)b1;
        else { u.begin[u.n] = (uint32_t)b0; u.end[u.n] = (uint32_t)b1; ++u.n; }
    }
    return u;
}

static hipError_t cache_adam(mrt_ctx* ctx, mrt_cache* c, uint32_t pos, uint32_t len) {
    c->step += 1;
    const float b1 = c->opts.adam_beta1, b2 = c->opts.adam_beta2;
    AdamHyper h;
    h.b1 = b1;
    h.b2 = b2;
    h.eps = c->opts.adam_eps;
    h.c1 = 1.0f - (float)std::pow((double)b1, (double)c->step);
    h.c2 = 1.0f - (float)std::pow((double)b2, (double)c->step);
    h.ema_a = c->opts.ema_decay;
    h.ema_aS = (float)std::pow((double)h.ema_a, (double)len);
    h.ema_gain = 1.0f - (float)std::pow((double)h.ema_a, (double)(len - pos));
    h.ema_base = pos == 0 ? 1u : 0u;
    h.ema_exact = len == 1 ? 1u : 0u;
    AdamStep a;
    a.W = c->d_W; a.gW = c->d_gW; a.mW = c->d_mW; a.vW = c->d_vW; a.eW = c->d_eW;
    a.T = c->d_table; a.gT = c->d_gtable; a.mT = c->d_mtable; a.vT = c->d_vtable; a.eT = c->d_etable;
    a.table_floats = c->table_floats;  // levels * 2^log2_table * 2 with log2_table >= 8: a multiple of 4
    a.lr_w = c->opts.lr;
    a.lr_t = c->opts.table_lr > 0.0f ? c->opts.table_lr : c->opts.lr;
    a.used = table_ranges(c);
    a.flag = c->d_flag + 1 + (c->step & 1u);
    a.flag_next = c->d_flag + 1 + ((c->step + 1u) & 1u);
    a.carry = pos == 0 ? 0u : 1u;
    c->step_flag_current = true;
    ProfScope p(ctx, "cache_adam");
    const size_t tblocks = std::min<size_t>(blocks_for(c->table_floats / 4), (size_t)ctx->num_cus * 16);
    hipLaunchKernelGGL(k_adam_step, dim3(kWeightBlocks + (uint32_t)tblocks), dim3(256), 0, ctx->stream, a, h);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    c->grad_clean = true;  // the step zeroed every gradient it consumed; the others were zero
    return cache_pack(ctx, c);
}

// ---- entry points shared with the path tracer (declared in internal.h) ----------------------
void cache_loss(const mrt_cache* c, const float** parts, uint32_t* blocks) {
    *parts = c->d_loss;
    *blocks = c->loss_blocks;
}

hipError_t cache_reserve(mrt_cache* c, size_t query_cap, size_t train_cap) {
    hipError_t e = scratch_ensure(c->inf, query_cap, false, c->grid);
    if (e == hipSuccess && train_cap) e = scratch_ensure(c->trn, train_cap, true, c->grid);
    return e;
}

hipError_t cache_query_batch(mrt_ctx* ctx, mrt_cache* c, const mrt_cache_sample* qs, const uint32_t* count,
                             uint32_t nmax, const float** y) {
    hipError_t e = scratch_ensure(c->inf, nmax, false, c->grid);
    if (e != hipSuccess) return e;
    *y = c->inf.y;
    if (nmax == 0) return hipSuccess;
    return cache_forward(ctx, c, c->inf, qs, count, nmax);
}

hipError_t cache_train_batch(mrt_ctx* ctx, mrt_cache* c, const mrt_cache_sample* qs, const float* target,
                             const uint32_t* count, uint32_t nmax, int apply, uint32_t group_pos, uint32_t group_len) {
    if (group_len == 0 || group_pos >= group_len) return hipErrorInvalidValue;
    if (nmax == 0) return apply ? cache_adam(ctx, c, group_pos, group_len) : hipSuccess;  // apply what is accumulated
    hipError_t e = scratch_ensure(c->trn, nmax, true, c->grid);
    if (e != hipSuccess) return e;
    if ((e = cache_train_pass(ctx, c, qs, target, count, nmax)) != hipSuccess) return e;
    return apply ? cache_adam(ctx, c, group_pos, group_len) : hipSuccess;
}

// Refreshes the non-finite flag from the current weights and table; returns its device address.
hipError_t cache_check(mrt_ctx* ctx, mrt_cache* c, const uint32_t** flag) {
    if (c->step_flag_current) {  // the last Adam step checked every parameter it left
        *flag = c->d_flag + 1 + (c->step & 1u);
        return hipSuccess;
    }
    hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(256), 0, ctx->stream, reinterpret_cast<float*>(c->d_flag), (size_t)1);
    hipLaunchKernelGGL(k_check_finite, dim3(blocks_for((size_t)kLayers * 4096)), dim3(256), 0, ctx->stream, c->d_W,
                       (size_t)kLayers * 4096, c->d_flag);
    const uint32_t tblocks = (uint32_t)std::min<size_t>(blocks_for(c->table_floats / 4), (size_t)ctx->num_cus * 16);
    hipLaunchKernelGGL(k_check_finite_ranges, dim3(tblocks), dim3(256), 0, ctx->stream,
                       reinterpret_cast<const float4*>(c->d_table), table_ranges(c), c->d_flag);
    if (ema_on(c)) {  // queries read the bias-corrected averages
        hipLaunchKernelGGL(k_check_finite, dim3(blocks_for((size_t)kLayers * 4096)), dim3(256), 0, ctx->stream, c->d_eW,
                           (size_t)kLayers * 4096, c->d_flag);
        hipLaunchKernelGGL(k_check_finite_ranges, dim3(tblocks), dim3(256), 0, ctx->stream,
                           reinterpret_cast<const float4*>(c->d_etable), table_ranges(c), c->d_flag);
    }
    *flag = c->d_flag;
    return hipGetLastError();
}

extern "C" {

mrt_status mrt_cache_create(mrt_ctx* ctx, const mrt_cache_opts* o, mrt_cache** out) {
    if (!ctx || !o || !out) return api_fail(ctx, MRT_ERR_INVALID, "mrt_cache_create: null argument");
    *out = nullptr;
    if (o->levels == 0 || o->levels > kMaxLevels) return api_fail(ctx, MRT_ERR_INVALID, "mrt_cache_create: levels 1..16");
    if (o->log2_table < 8 || o->log2_table > 24) return api_fail(ctx, MRT_ERR_INVALID, "mrt_cache_create: log2_table 8..24");
    if (!(o->ema_decay >= 0.0f && o->ema_decay < 1.0f))
        return api_fail(ctx, MRT_ERR_INVALID, "mrt_cache_create: ema_decay must be in [0, 1)");
    if (!(o->table_lr >= 0.0f && o->table_lr < 1.0f))
        return api_fail(ctx, MRT_ERR_INVALID, "mrt_cache_create: table_lr must be in [0, 1)");
    if (!(o->base_res >= 1.0f) || !(o->max_res >= o->base_res))
        return api_fail(ctx, MRT_ERR_INVALID, "mrt_cache_create: need 1 <= base_res <= max_res");
    for (int k = 0; k < 3; ++k)
        if (!(o->hi[k] > o->lo[k])) return api_fail(ctx, MRT_ERR_INVALID, "mrt_cache_create: empty bounds");
    auto* c = new (std::nothrow) mrt_cache();
    if (!c) return api_fail(ctx, MRT_ERR_OOM, "mrt_cache_create: host allocation failed");
    c->opts = *o;
    GridDev& g = c->grid;
    g.levels = o->levels;
    g.log2_table = o->log2_table;
    const double b = o->levels > 1 ? std::exp((std::log((double)o->max_res) - std::log((double)o->base_res)) /
                                              (double)(o->levels - 1))
                                   : 1.0;
    const uint64_t tsize = 1ull << o->log2_table;
    for (uint32_t l = 0; l < o->levels; ++l) {
        g.res[l] = (float)std::floor((double)o->base_res * std::pow(b, (double)l));
        const uint64_t side = (uint64_t)g.res[l] + 1;
        g.dense[l] = side * side * side <= tsize ? 1u : 0u;
    }
    for (int k = 0; k < 3; ++k) { g.lo[k] = o->lo[k]; g.inv_ext[k] = 1.0f / (o->hi[k] - o->lo[k]); }
    c->table_floats = (size_t)o->levels * tsize * 2;
    hipError_t e = hipSuccess;
    auto A = [&](auto** p, size_t bytes) {
        if (e == hipSuccess) { e = hipMalloc((void**)p, bytes); c->device_bytes += bytes; }
    };
    const size_t wbytes = (size_t)kLayers * 4096 * 4;
    A(&c->d_W, wbytes); A(&c->d_gW, wbytes); A(&c->d_mW, wbytes); A(&c->d_vW, wbytes);
    A(&c->d_table, c->table_floats * 4); A(&c->d_gtable, c->table_floats * 4);
    A(&c->d_mtable, c->table_floats * 4); A(&c->d_vtable, c->table_floats * 4);
    A(&c->d_wf, (size_t)kLayers * kLayerPack); A(&c->d_wt, (size_t)kLayers * kLayerPack);
    A(&c->d_wq, (size_t)kLayers * kLayerPack);
    A(&c->d_loss, (kMaxLossBlocks + 1) * 4); A(&c->d_flag, 256);
    if (o->ema_decay > 0.0f) {
        A(&c->d_eW, wbytes); A(&c->d_etable, c->table_floats * 4);
    }
    if (e == hipSuccess) {
        // Xavier-uniform weights (rows >= 3 of the output layer zero), table U(-1e-4, 1e-4).
        std::mt19937 rng(o->seed);
        std::vector<float> w((size_t)kLayers * 4096, 0.0f);
        const float a = std::sqrt(6.0f / (float)(kW + kW));
        std::uniform_real_distribution<float> uw(-a, a), ut(-1e-4f, 1e-4f);
        for (int L = 0; L < kLayers; ++L)
            for (int r = 0; r < kW; ++r)
                for (int k = 0; k < kW; ++k) w[(size_t)L * 4096 + r * kW + k] = (L == kLayers - 1 && r >= 3) ? 0.0f : uw(rng);
        std::vector<float> t(c->table_floats);
        for (auto& x : t) x = ut(rng);
        e = hipMemcpy(c->d_W, w.data(), wbytes, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(c->d_table, t.data(), c->table_floats * 4, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(256), 0, ctx->stream, reinterpret_cast<float*>(c->d_flag), (size_t)64);
        hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(256), 0, ctx->stream, c->d_loss, (size_t)kMaxLossBlocks + 1);
        hipLaunchKernelGGL(k_zero_f32, dim3(64), dim3(256), 0, ctx->stream, c->d_gW, (size_t)kLayers * 4096);
        hipLaunchKernelGGL(k_zero_f32, dim3(64), dim3(256), 0, ctx->stream, c->d_mW, (size_t)kLayers * 4096);
        hipLaunchKernelGGL(k_zero_f32, dim3(64), dim3(256), 0, ctx->stream, c->d_vW, (size_t)kLayers * 4096);
        hipLaunchKernelGGL(k_zero_f32, dim3(1024), dim3(256), 0, ctx->stream, c->d_gtable, c->table_floats);
        hipLaunchKernelGGL(k_zero_f32, dim3(1024), dim3(256), 0, ctx->stream, c->d_mtable, c->table_floats);
        hipLaunchKernelGGL(k_zero_f32, dim3(1024), dim3(256), 0, ctx->stream, c->d_vtable, c->table_floats);
        if (c->d_eW) {
            hipLaunchKernelGGL(k_zero_f32, dim3(64), dim3(256), 0, ctx->stream, c->d_eW, (size_t)kLayers * 4096);
            hipLaunchKernelGGL(k_zero_f32, dim3(1024), dim3(256), 0, ctx->stream, c->d_etable, c->table_floats);
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = cache_pack(ctx, c);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        cache_free(c);
        return api_fail(ctx, e == hipErrorOutOfMemory ? MRT_ERR_OOM : MRT_ERR_HIP, "mrt_cache_create: %s",
                        hipGetErrorString(e));
    }
    *out = c;
    return MRT_OK;
}

mrt_status mrt_cache_destroy(mrt_ctx* ctx, mrt_cache* c) {
    if (!ctx) return api_fail(nullptr, MRT_ERR_INVALID, "mrt_cache_destroy: null context");
    if (!c) return MRT_OK;
    drop_training(ctx, c);  // deferred steps of a cache nobody can read any more
    hipStreamSynchronize(ctx->stream);
    hipStreamSynchronize(ctx->side);
    cache_free(c);
    return MRT_OK;
}

mrt_status mrt_cache_query(mrt_ctx* ctx, mrt_cache* c, const mrt_cache_sample* q, uint32_t n, float* rgb4) {
    if (!ctx || !c) return api_fail(ctx, MRT_ERR_INVALID, "mrt_cache_query: null context or cache");
    if (n && (!q || !rgb4)) return api_fail(ctx, MRT_ERR_INVALID, "mrt_cache_query: null buffer");
    const float* y = nullptr;
    hipError_t e = flush_training(ctx, c);
    if (e == hipSuccess) e = cache_query_batch(ctx, c, q, nullptr, n, &y);
    if (e == hipSuccess && n) e = hipMemcpyAsync(rgb4, y, (size_t)n * 16, hipMemcpyDeviceToDevice, ctx->stream);
    if (e != hipSuccess) return api_fail(ctx, MRT_ERR_HIP, "mrt_cache_query: %s", hipGetErrorString(e));
    return MRT_OK;
}

mrt_status mrt_cache_inputs(mrt_ctx* ctx, mrt_cache* c, const mrt_cache_sample* q, uint32_t n, float* x64) {
    if (!ctx || !c) return api_fail(ctx, MRT_ERR_INVALID, "mrt_cache_inputs: null context or cache");
    if (n && (!q || !x64)) return api_fail(ctx, MRT_ERR_INVALID, "mrt_cache_inputs: null buffer");
    if (n == 0) return MRT_OK;
    hipError_t e = flush_training(ctx, c);
    if (e == hipSuccess) e = scratch_ensure(c->inf, n, false, c->grid);
    if (e == hipSuccess) {
        launch_encode(ctx, c, c->d_table, 1.0f, q, nullptr, n, (uint32_t)c->inf.cap, c->inf.x0, nullptr);
        hipLaunchKernelGGL(k_bf16_to_f32, dim3(blocks_for((size_t)n * kW)), dim3(256), 0, ctx->stream, c->inf.x0, x64,
                           (size_t)n * kW);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return api_fail(ctx, MRT_ERR_HIP, "mrt_cache_inputs: %s", hipGetErrorString(e));
    return MRT_OK;
}

mrt_status mrt_cache_train(mrt_ctx* ctx, mrt_cache* c, const mrt_cache_sample* q, const float* target4, uint32_t n,
                           int apply) {
    if (!ctx || !c) return api_fail(ctx, MRT_ERR_INVALID, "mrt_cache_train: null context or cache");
    if (n && (!q || !target4)) return api_fail(ctx, MRT_ERR_INVALID, "mrt_cache_train: null buffer");
    hipError_t e = flush_training(ctx, c);
    if (e == hipSuccess) e = cache_train_batch(ctx, c, q, target4, nullptr, n, apply);
    if (e != hipSuccess) return api_fail(ctx, MRT_ERR_HIP, "mrt_cache_train: %s", hipGetErrorString(e));
    return MRT_OK;
}

mrt_status mrt_cache_stats(mrt_ctx* ctx, mrt_cache* c, float* last_loss_sum, uint32_t* steps, uint32_t* nonfinite) {
    if (!ctx || !c) return api_fail(ctx, MRT_ERR_INVALID, "mrt_cache_stats: null argument");
    hipError_t e = flush_training(ctx, c);
    if (e == hipSuccess && nonfinite) {
        const uint32_t* flag = nullptr;
        e = cache_check(ctx, c, &flag);
        if (e == hipSuccess) e = hipMemcpyAsync(nonfinite, flag, 4, hipMemcpyDeviceToHost, ctx->stream);
    }
    float part[kMaxLossBlocks];
    if (e == hipSuccess && last_loss_sum && c->loss_blocks)
        e = hipMemcpyAsync(part, c->d_loss, c->loss_blocks * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return api_fail(ctx, MRT_ERR_HIP, "mrt_cache_stats: %s", hipGetErrorString(e));
    if (last_loss_sum) {
        double sum = 0.0;
        for (uint32_t b = 0; b < c->loss_blocks; ++b) sum += part[b];
        *last_loss_sum = (float)sum;
    }
    if (steps) *steps = c->step;
    return MRT_OK;
}

// which: 0 = MLP weights (6*64*64), 1 = hash table (levels*2^T*2), 2 = MLP weight gradient,
// 3 = table gradient, 4 / 5 = running averages of the weights / table (ema_decay > 0).
// Host buffers; `count` floats must match.
mrt_status mrt_cache_get(mrt_ctx* ctx, mrt_cache* c, uint32_t which, float* host, size_t count) {
    if (!ctx || !c || !host) return api_fail(ctx, MRT_ERR_INVALID, "mrt_cache_get: null argument");
    const float* src = which == 0 ? c->d_W : which == 1 ? c->d_table : which == 2 ? c->d_gW : which == 3 ? c->d_gtable
                     : which == 4 ? c->d_eW : which == 5 ? c->d_etable : nullptr;
    const size_t expect = (which == 0 || which == 2 || which == 4) ? (size_t)kLayers * 4096 : c->table_floats;
    if (!src || count != expect) return api_fail(ctx, MRT_ERR_INVALID, "mrt_cache_get: which %u count %zu (want %zu)",
                                                 which, count, expect);
    hipError_t e = flush_training(ctx, c);
    if (e == hipSuccess) e = hipMemcpyAsync(host, src, count * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return api_fail(ctx, MRT_ERR_HIP, "mrt_cache_get: %s", hipGetErrorString(e));
    return MRT_OK;
}

mrt_status mrt_cache_set(mrt_ctx* ctx, mrt_cache* c, uint32_t which, const float* host, size_t count) {
    if (!ctx || !c || !host) return api_fail(ctx, MRT_ERR_INVALID, "mrt_cache_set: null argument");
    float* dst = which == 0 ? c->d_W : which == 1 ? c->d_table : nullptr;
    const size_t expect = which == 0 ? (size_t)kLayers * 4096 : c->table_floats;
    if (!dst || count != expect) return api_fail(ctx, MRT_ERR_INVALID, "mrt_cache_set: which %u count %zu (want %zu)",
                                                 which, count, expect);
    hipError_t e = flush_training(ctx, c);
    if (e == hipSuccess) e = hipMemcpyAsync(dst, host, count * 4, hipMemcpyHostToDevice, ctx->stream);
    float* avg = which == 0 ? c->d_eW : c->d_etable;
    if (e == hipSuccess && avg) {
        // The running average restarts at the set values: e = p (1 - a^steps), whose bias-corrected
        // value e / (1 - a^steps) is p, so queries see the parameters just set (and no earlier
        // history, finite or not, survives a host update).
        const float keep = (float)(1.0 - std::pow((double)c->opts.ema_decay, (double)c->step));
        hipLaunchKernelGGL(k_scale_copy, dim3((uint32_t)std::min<size_t>(blocks_for(count), (size_t)ctx->num_cus * 16)),
                           dim3(256), 0, ctx->stream, avg, dst, count, keep);
        e = hipGetLastError();
    }
    if (e == hipSuccess && which == 0) e = cache_pack(ctx, c);
    c->step_flag_current = false;  // parameters changed outside an Adam step
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return api_fail(ctx, MRT_ERR_HIP, "mrt_cache_set: %s", hipGetErrorString(e));
    return MRT_OK;
}

}  // extern "C"
