// oracle_transc.hpp — the oracle's restatement of exp, sin, cos, acos and x^1.5 from IEEE basic
// operations (test infrastructure only). Same methods as the library uses for the sky:
//   exp: k = round(x / ln 2); r = (x - k*ln2_hi) - k*ln2_lo (ln2_hi with 12 zero low bits);
//        Taylor degree 7; result ldexp(p, k).
//   sin/cos: n = round(2x/pi); y = ((x - n*c1) - n*c2) - n*c3, c1+c2+c3 = pi/2; Taylor 9 / 10;
//        quadrant n mod 4.
//   acos: asin(z) = z + z^3 P(z^2) on |z| <= 1/2 with P a degree-5 least-squares fit; reflected
//        through 2 asin(sqrt((1-|x|)/2)) outside.
//   x^1.5 = x sqrt(x).
//   log: frexp into [sqrt(1/2), sqrt(2)), ln m = 2 atanh((m-1)/(m+1)) to s^11; pow = exp(y log x).
// Pinned against libm in double by tests/test_oracle_sky.py.
#pragma once
#include <cmath>

namespace orc_tx {

inline float exp(float x) {
    if (x > 88.7f) return HUGE_VALF;
    if (x < -104.0f) return 0.0f;
    const float k = std::floor(x * 1.44269502f + 0.5f);
    const float r = (x - k * 0.693115234375f) - k * 3.19461833e-05f;
    const float c[8] = {1.0f / 5040.0f, 1.0f / 720.0f, 1.0f / 120.0f, 1.0f / 24.0f, 1.0f / 6.0f, 0.5f, 1.0f, 1.0f};
    float p = c[0];
    for (int i = 1; i < 8; ++i) p = p * r + c[i];
    return std::ldexp(p, (int)k);
}

inline void sincos(float x, float& s, float& c) {
    const float n = std::floor(x * 0.636619747f + 0.5f);
    const float y = ((x - n * 1.5703125f) - n * 4.83751297e-04f) - n * 7.54979013e-08f;
    const float y2 = y * y;
    const float cs[4] = {1.0f / 362880.0f, -1.0f / 5040.0f, 1.0f / 120.0f, -1.0f / 6.0f};
    float ps = cs[0];
    for (int i = 1; i < 4; ++i) ps = ps * y2 + cs[i];
    const float sy = y + (y * y2) * ps;
    const float cc[5] = {-1.0f / 3628800.0f, 1.0f / 40320.0f, -1.0f / 720.0f, 1.0f / 24.0f, -0.5f};
    float pc = cc[0];
    for (int i = 1; i < 5; ++i) pc = pc * y2 + cc[i];
    const float cy = 1.0f + y2 * pc;
    switch (((int)n) & 3) {
        case 0: s = sy; c = cy; break;
        case 1: s = cy; c = -sy; break;
        case 2: s = -sy; c = -cy; break;
        default: s = -cy; c = sy; break;
    }
}
inline float sin(float x) { float s, c; sincos(x, s, c); return s; }
inline float cos(float x) { float s, c; sincos(x, s, c); return c; }

inline float asin_small(float z) {
    const float w = z * z;
    const float cf[6] = {0.0374410264f, 0.0144919502f, 0.0317939259f, 0.0445185229f, 0.0750049725f, 0.166666597f};
    float p = cf[0];
    for (int i = 1; i < 6; ++i) p = p * w + cf[i];
    return z + (z * w) * p;
}

inline float acos(float x) {
    x = std::fmin(std::fmax(x, -1.0f), 1.0f);
    if (std::fabs(x) <= 0.5f) return 1.57079637f - asin_small(x);
    const float z = std::sqrt((1.0f - std::fabs(x)) * 0.5f);
    const float a = 2.0f * asin_small(z);
    return x > 0.0f ? a : 3.14159274f - a;
}

inline float pow15(float x) { return x * std::sqrt(x); }

inline float log(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -HUGE_VALF : NAN;
    if (x == HUGE_VALF) return x;
    int e;
    float m = std::frexp(x, &e);
    if (m < 0.707106781f) { m = m * 2.0f; e -= 1; }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    const float c[6] = {1.0f / 11.0f, 1.0f / 9.0f, 1.0f / 7.0f, 1.0f / 5.0f, 1.0f / 3.0f, 1.0f};
    float p = c[0];
    for (int i = 1; i < 6; ++i) p = p * s2 + c[i];
    const float lm = (2.0f * s) * p;
    const float fe = (float)e;
    return fe * 0.693115234375f + (lm + fe * 3.19461833e-05f);
}

inline float pow(float x, float y) {
    if (x == 0.0f) return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : HUGE_VALF);
    return exp(y * log(x));
}

}  // namespace orc_tx
