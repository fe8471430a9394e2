// oracle_cache.cpp — scalar CPU restatement of the radiance cache: hash-grid encoding and its
// gradient, MLP forward/backward with BF16 operands and FP32 accumulation, relative L2 loss,
// Adam. TEST INFRASTRUCTURE ONLY (see oracle_cache.hpp for the published methods).
#include "oracle.h"
#include "oracle_cache.hpp"

#include <vector>

using namespace orc;

namespace {

GridParams make_grid(const orc_grid* g) {
    GridParams p;
    p.levels = g->levels;
    p.log2_table = g->log2_table;
    p.base_res = g->base_res;
    p.max_res = g->max_res;
    for (int k = 0; k < 3; ++k) {
        p.lo[k] = g->lo[k];
        p.inv_ext[k] = 1.0f / (g->hi[k] - g->lo[k]);
    }
    return p;
}

}  // namespace

extern "C" {

float orc_bf16_round(float f) { return bf16_round(f); }

void orc_cache_encode(const orc_grid* gp, const float* table, const float* pos, const float* dir, const float* nrm,
                      const float* alb, uint32_t n, float scale, float* x0) {
    GridParams g = make_grid(gp);
    const size_t tsize = (size_t)1 << g.log2_table;
    for (uint32_t q = 0; q < n; ++q) {
        float* x = x0 + (size_t)q * kMlpWidth;
        for (int k = 0; k < kMlpWidth; ++k) x[k] = 0.0f;
        for (uint32_t l = 0; l < g.levels; ++l) {
            uint32_t idx[8];
            float w[8];
            grid_corners(g, l, level_scale(g, l), pos + 3 * (size_t)q, idx, w);
            float f0 = 0.0f, f1 = 0.0f;
            for (int c = 0; c < 8; ++c) {
                const float* t = table + ((size_t)l * tsize + idx[c]) * 2;
                f0 = f0 + w[c] * t[0];
                f1 = f1 + w[c] * t[1];
            }
            x[2 * l] = bf16_round(f0 * scale);
            x[2 * l + 1] = bf16_round(f1 * scale);
        }
        for (int k = 0; k < 3; ++k) {
            x[kEncDims + k] = bf16_round(dir[3 * (size_t)q + k]);
            x[kEncDims + 3 + k] = bf16_round(nrm[3 * (size_t)q + k]);
            x[kEncDims + 6 + k] = bf16_round(alb[3 * (size_t)q + k]);
        }
        for (int k = kEncDims + 9; k < kMlpWidth; ++k) x[k] = 0.0f;
    }
}

void orc_mlp_forward(const float* W, const float* x0, uint32_t n, float* y, float* acts) {
    std::vector<float> wb((size_t)kMlpLayers * kMlpWidth * kMlpWidth);
    for (size_t i = 0; i < wb.size(); ++i) wb[i] = bf16_round(W[i]);
    float a[kMlpWidth], z[kMlpWidth];
    for (uint32_t q = 0; q < n; ++q) {
        std::memcpy(a, x0 + (size_t)q * kMlpWidth, sizeof(a));
        for (int L = 0; L < kMlpLayers; ++L) {
            const float* w = &wb[(size_t)L * kMlpWidth * kMlpWidth];
            for (int o = 0; o < kMlpWidth; ++o) {
                float s = 0.0f;
                for (int k = 0; k < kMlpWidth; ++k) s = s + w[o * kMlpWidth + k] * a[k];
                z[o] = s;
            }
            if (L + 1 < kMlpLayers) {
                for (int o = 0; o < kMlpWidth; ++o) a[o] = bf16_round(std::fmax(z[o], 0.0f));
                if (acts) std::memcpy(acts + ((size_t)L * n + q) * kMlpWidth, a, sizeof(a));
            }
        }
        for (int c = 0; c < 4; ++c) y[(size_t)q * 4 + c] = c < 3 ? z[c] : 0.0f;
    }
}

// Returns the mean relative L2 loss; dW (6*64*64) and dx0 (n*64) are overwritten.
double orc_mlp_backward(const float* W, const float* x0, uint32_t n, const float* target, float* dW, float* dx0) {
    std::vector<float> wb((size_t)kMlpLayers * kMlpWidth * kMlpWidth);
    for (size_t i = 0; i < wb.size(); ++i) wb[i] = bf16_round(W[i]);
    std::vector<float> acts((size_t)(kMlpLayers - 1) * n * kMlpWidth), y((size_t)n * 4);
    orc_mlp_forward(W, x0, n, y.data(), acts.data());
    std::vector<float> dnext((size_t)n * kMlpWidth, 0.0f), dcur((size_t)n * kMlpWidth, 0.0f);
    double loss = 0.0;
    const float inv_n = 1.0f / (float)n;
    for (uint32_t q = 0; q < n; ++q) {
        for (int c = 0; c < 3; ++c) {
            const float yc = y[(size_t)q * 4 + c], tc = target[(size_t)q * 4 + c];
            const float den = yc * yc + 0.01f;
            const float diff = yc - tc;
            loss += (double)(diff * diff / den);
            dnext[(size_t)q * kMlpWidth + c] = bf16_round(((2.0f * diff) / den) * inv_n);
        }
    }
    loss /= (double)n;
    std::memset(dW, 0, sizeof(float) * kMlpLayers * kMlpWidth * kMlpWidth);
    for (int L = kMlpLayers - 1; L >= 0; --L) {
        const float* ain = L > 0 ? &acts[(size_t)(L - 1) * n * kMlpWidth] : x0;
        float* dw = dW + (size_t)L * kMlpWidth * kMlpWidth;
        for (uint32_t q = 0; q < n; ++q) {
            const float* d = &dnext[(size_t)q * kMlpWidth];
            const float* a = ain + (size_t)q * kMlpWidth;
            for (int o = 0; o < kMlpWidth; ++o) {
                if (d[o] == 0.0f) continue;
                for (int i = 0; i < kMlpWidth; ++i) dw[o * kMlpWidth + i] = dw[o * kMlpWidth + i] + d[o] * a[i];
            }
        }
        const float* w = &wb[(size_t)L * kMlpWidth * kMlpWidth];
        for (uint32_t q = 0; q < n; ++q) {
            const float* d = &dnext[(size_t)q * kMlpWidth];
            const float* a = ain + (size_t)q * kMlpWidth;
            for (int i = 0; i < kMlpWidth; ++i) {
                float s = 0.0f;
                for (int o = 0; o < kMlpWidth; ++o) s = s + w[o * kMlpWidth + i] * d[o];
                if (L > 0) dcur[(size_t)q * kMlpWidth + i] = bf16_round(a[i] > 0.0f ? s : 0.0f);
                else dx0[(size_t)q * kMlpWidth + i] = s;
            }
        }
        dnext.swap(dcur);
    }
    return loss;
}

void orc_cache_encode_backward(const orc_grid* gp, const float* pos, const float* dx0, uint32_t n, float* grad_table) {
    GridParams g = make_grid(gp);
    const size_t tsize = (size_t)1 << g.log2_table;
    for (uint32_t q = 0; q < n; ++q) {
        for (uint32_t l = 0; l < g.levels; ++l) {
            uint32_t idx[8];
            float w[8];
            grid_corners(g, l, level_scale(g, l), pos + 3 * (size_t)q, idx, w);
            const float g0 = dx0[(size_t)q * kMlpWidth + 2 * l], g1 = dx0[(size_t)q * kMlpWidth + 2 * l + 1];
            for (int c = 0; c < 8; ++c) {
                float* t = grad_table + ((size_t)l * tsize + idx[c]) * 2;
                t[0] += w[c] * g0;
                t[1] += w[c] * g1;
            }
        }
    }
}

void orc_adam(float* p, const float* g, float* m, float* v, uint32_t n, uint32_t step, float lr, float b1, float b2,
              float eps, int sparse) {
    const float c1 = 1.0f - (float)std::pow((double)b1, (double)step);
    const float c2 = 1.0f - (float)std::pow((double)b2, (double)step);
    for (uint32_t i = 0; i < n; ++i) {
        if (sparse && g[i] == 0.0f) continue;  // untouched hash entries keep parameter and moments
        m[i] = b1 * m[i] + (1.0f - b1) * g[i];
        v[i] = b2 * v[i] + (1.0f - b2) * (g[i] * g[i]);
        const float mh = m[i] / c1, vh = v[i] / c2;
        p[i] = p[i] - lr * mh / (std::sqrt(vh) + eps);
    }
}

}  // extern "C"
