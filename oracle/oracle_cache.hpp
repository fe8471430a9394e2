// oracle_cache.hpp — scalar CPU restatement of the radiance cache (test infrastructure only).
//
// No reference counterpart: the reference has no learned component. Published methods:
//   * multiresolution hash encoding (Mueller, Evans, Schied, Keller, "Instant Neural Graphics
//     Primitives", SIGGRAPH 2022): L levels, F features per level, table of 2^T entries per
//     level, resolution N_l = floor(N_min * b^l), b = exp((ln N_max - ln N_min) / (L - 1)),
//     dense indexing while (N_l + 1)^3 <= 2^T, else the spatial hash
//     (x * 1) ^ (y * 2654435761) ^ (z * 805459861) mod 2^T, trilinear interpolation;
//   * an MLP of six 64x64 layers (ReLU after the first five, linear output; the last layer's
//     rows >= 3 are padding and stay zero), BF16 operands with FP32 accumulation;
//   * relative L2 loss (Mueller et al., "Real-time Neural Radiance Caching", 2021):
//     sum_c (y_c - t_c)^2 / (sg(y_c)^2 + 0.01), averaged over samples;
//   * Adam (Kingma & Ba 2015) with bias correction.
// BF16 rounding is round-to-nearest-even of the FP32 bit pattern.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

constexpr int kMlpWidth = 64;
constexpr int kMlpLayers = 6;
constexpr int kEncDims = 32;

static inline float bf16_round(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) {  // inf / nan: keep, quiet the nan
        if (u & 0x007FFFFFu) u |= 0x00400000u;
        u &= 0xFFFF0000u;
    } else {
        u += 0x7FFFu + ((u >> 16) & 1u);
        u &= 0xFFFF0000u;
    }
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

struct GridParams {
    uint32_t levels, log2_table;
    float base_res, max_res;
    float lo[3], inv_ext[3];
};

static inline float level_scale(const GridParams& g, uint32_t l) {
    const double b = std::exp((std::log((double)g.max_res) - std::log((double)g.base_res)) / (double)(g.levels - 1));
    return (float)std::floor((double)g.base_res * std::pow(b, (double)l));
}

static inline uint32_t grid_index(uint32_t x, uint32_t y, uint32_t z, uint32_t res, uint32_t tsize) {
    const uint64_t side = (uint64_t)res + 1;
    if (side * side * side <= tsize) return (uint32_t)(x + side * (y + side * z));
    return (x * 1u ^ y * 2654435761u ^ z * 805459861u) & (tsize - 1u);
}

// Corner indices and trilinear weights of one level (corner c: bit0 x, bit1 y, bit2 z).
static inline void grid_corners(const GridParams& g, uint32_t l, float res_f, const float p[3], uint32_t idx[8],
                                float w[8]) {
    const uint32_t res = (uint32_t)res_f;
    const uint32_t tsize = 1u << g.log2_table;
    float fr[3];
    uint32_t c0[3];
    for (int k = 0; k < 3; ++k) {
        float u = (p[k] - g.lo[k]) * g.inv_ext[k];
        u = std::fmin(std::fmax(u, 0.0f), 1.0f);
        const float x = u * res_f;
        float fl = std::floor(x);
        if (fl >= res_f) fl = res_f - 1.0f;
        c0[k] = (uint32_t)fl;
        fr[k] = x - fl;
    }
    for (int c = 0; c < 8; ++c) {
        const uint32_t dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
        const float wx = dx ? fr[0] : 1.0f - fr[0];
        const float wy = dy ? fr[1] : 1.0f - fr[1];
        const float wz = dz ? fr[2] : 1.0f - fr[2];
        w[c] = (wx * wy) * wz;
        idx[c] = grid_index(c0[0] + dx, c0[1] + dy, c0[2] + dz, res, tsize);
    }
    (void)l;
}

}  // namespace orc
