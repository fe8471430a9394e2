// transc.h — exp, sin, cos, acos and x^1.5 from IEEE +, -, *, /, sqrt, floor and ldexp only, so
// that host and device evaluate them to the same bits (the sky's float positions sit at planetary
// scale, where one ulp of a direction or density moves an altitude |P| - R by 0.5 m and the
// Mie density by 0.04%; identical functions keep the GPU LUTs equal to the CPU restatement).
//   exp:      k = round(x / ln 2), Cody-Waite reduction with ln 2 split in two floats (the
//             first with 12 zero low bits, so k * hi is exact), Taylor degree 7 on |r| <= ln2/2
//             (truncation 5e-9 relative), scaled by ldexp.
//   sin, cos: n = round(x * 2/pi), Cody-Waite with pi/2 in three floats (exact n * hi for the
//             |x| < 16 used here), Taylor degree 9 / 10 on |y| <= pi/4, quadrant by n mod 4.
//   acos:     asin(z) = z + z^3 P(z^2) on |z| <= 1/2 with P a degree-5 least-squares fit (relative
//             error 0.02 ulp before rounding), acos(x) = pi/2 - asin(x) for |x| <= 1/2, else
//             2 asin(sqrt((1 - |x|) / 2)) (reflected for x < 0).
//   pow15:    x^1.5 = x * sqrt(x) for x >= 0.
//   log, pow: frexp into [sqrt(1/2), sqrt(2)), 2 atanh series; pow(x, y) = exp(y log x), x >= 0.
// Every operation is separately rounded (-ffp-contract=off on both sides).
#pragma once
#include "rt_math.h"

namespace rt {

RT_HD float tx_exp(float x) {
    if (x > 88.7f) return __builtin_huge_valf();
    if (x < -104.0f) return 0.0f;
    const float k = floorf(x * 1.44269502f + 0.5f);
    const float r = (x - k * 0.693115234375f) - k * 3.19461833e-05f;
    float p = 1.0f / 5040.0f;
    p = p * r + 1.0f / 720.0f;
    p = p * r + 1.0f / 120.0f;
    p = p * r + 1.0f / 24.0f;
    p = p * r + 1.0f / 6.0f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    return ldexpf(p, (int)k);
}

RT_HD void tx_sincos(float x, float& s, float& c) {
    const float n = floorf(x * 0.636619747f + 0.5f);
    const float y = ((x - n * 1.5703125f) - n * 4.83751297e-04f) - n * 7.54979013e-08f;
    const float y2 = y * y;
    float ps = 1.0f / 362880.0f;
    ps = ps * y2 - 1.0f / 5040.0f;
    ps = ps * y2 + 1.0f / 120.0f;
    ps = ps * y2 - 1.0f / 6.0f;
    const float sy = y + (y * y2) * ps;
    float pc = -1.0f / 3628800.0f;
    pc = pc * y2 + 1.0f / 40320.0f;
    pc = pc * y2 - 1.0f / 720.0f;
    pc = pc * y2 + 1.0f / 24.0f;
    pc = pc * y2 - 0.5f;
    const float cy = 1.0f + y2 * pc;
    const int q = ((int)n) & 3;
    s = q == 0 ? sy : q == 1 ? cy : q == 2 ? -sy : -cy;
    c = q == 0 ? cy : q == 1 ? -sy : q == 2 ? -cy : sy;
}
RT_HD float tx_sin(float x) { float s, c; tx_sincos(x, s, c); return s; }
RT_HD float tx_cos(float x) { float s, c; tx_sincos(x, s, c); return c; }

RT_HD float tx_asin_small(float z) {  // |z| <= 1/2
    const float w = z * z;
    float p = 0.0374410264f;
    p = p * w + 0.0144919502f;
    p = p * w + 0.0317939259f;
    p = p * w + 0.0445185229f;
    p = p * w + 0.0750049725f;
    p = p * w + 0.166666597f;
    return z + (z * w) * p;
}

RT_HD float tx_acos(float x) {
    x = fminf(fmaxf(x, -1.0f), 1.0f);
    if (fabsf(x) <= 0.5f) return 1.57079637f - tx_asin_small(x);
    const float z = sqrtf((1.0f - fabsf(x)) * 0.5f);
    const float a = 2.0f * tx_asin_small(z);
    return x > 0.0f ? a : 3.14159274f - a;
}

RT_HD float tx_pow15(float x) { return x * sqrtf(x); }

// log: x = m 2^e with m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s), s = (m-1)/(m+1), |s| <= 0.1716,
// series to s^11 (truncation 1e-10 relative); ln x = e ln2_hi + (ln m + e ln2_lo).
RT_HD float tx_log(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -__builtin_huge_valf() : __builtin_nanf("");
    if (x == __builtin_huge_valf()) return x;
    int e;
    float m = frexpf(x, &e);
    if (m < 0.707106781f) { m = m * 2.0f; e -= 1; }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float p = 1.0f / 11.0f;
    p = p * s2 + 1.0f / 9.0f;
    p = p * s2 + 1.0f / 7.0f;
    p = p * s2 + 1.0f / 5.0f;
    p = p * s2 + 1.0f / 3.0f;
    p = p * s2 + 1.0f;
    const float lm = (2.0f * s) * p;
    const float fe = (float)e;
    return fe * 0.693115234375f + (lm + fe * 3.19461833e-05f);
}

// pow for x >= 0 (the tone curves only raise non-negative values): exp(y ln x)
RT_HD float tx_pow(float x, float y) {
    if (x == 0.0f) return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : __builtin_huge_valf());
    return tx_exp(y * tx_log(x));
}

}  // namespace rt
