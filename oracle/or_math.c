/*
 * or_math.c -- Philox4x32-10 and the FP32 function definitions of the oracle.
 * TEST INFRASTRUCTURE ONLY (see emtf_oracle.h).
 *
 * Reference call sites being restated:
 *   - RNG: curand_init / curand / curand_uniform, proposed model/SG-EMTF.cu:163-179
 *     (uses at :151, :187-188, :572-576, :717-719, :806, :873, :1552, :1597).
 *     cuRAND XORWOW is a third-party stream that cannot be reproduced here; it
 *     is replaced by Philox4x32-10 (parity of the stream itself: unpinned by
 *     design; the Philox implementation is pinned by known-answer vectors).
 *   - libm: cospif / sinf / expf / sqrtf / powf in SG-EMTF.cu:548-606, :711-749,
 *     :898-1111 and SG-EMTF_DE.cu:834-1069.  Replaced by the definitions below
 *     (Cody-Waite style reduction + Chebyshev polynomials, coefficients from
 *     oracle/fit_coeffs.py), pinned against float64 libm in the tests.
 */
#include "emtf_oracle.h"
#include <math.h>
#include <string.h>

#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void or_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += PHILOX_W0; k1 += PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void or_draw_block(uint64_t seed, uint32_t task, uint32_t gen, uint32_t purpose,
                   uint32_t index, uint32_t sub, uint32_t block, uint32_t out[4]) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t ctr[4] = {block, (purpose << 24) | ((sub & 0xFFu) << 16) | (index & 0xFFFFu), gen, task};
    or_philox(ctr, key, out);
}

/* ---- helpers ------------------------------------------------------------ */
static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
/* 2^n for -126 <= n <= 127, exact. */
static inline float pow2i(int n) { return u2f((uint32_t)(n + 127) << 23); }

/* sin(pi a) = a * SP(a^2), a in [0, 1/2] */
static const float SPC[5] = {7.765591145e-02f, -5.982903838e-01f, 2.550077438e+00f, -5.167710304e+00f,
                             3.141592741e+00f};
/* sin(2 pi b) = b * SN(x), x = b^2, b in [0, 1/4] */
static const float SNC[7] = {3.665855169e+00f, -1.508148289e+01f, 4.205813599e+01f, -7.670584869e+01f,
                             8.160524750e+01f, -4.134170151e+01f, 6.283185482e+00f};
/* expm1(r) = r + r^2 E(r), |r| <= ln2/2 */
static const float EC[7] = {2.485957884e-05f, 1.989927405e-04f, 1.388885430e-03f, 8.333298378e-03f,
                            4.166666791e-02f, 1.666666716e-01f, 5.000000000e-01f};
/* 2^f, |f| <= 1/2 */
static const float X2C[8] = {1.530369991e-05f, 1.546972926e-04f, 1.333347871e-03f, 9.618025273e-03f,
                             5.550410971e-02f, 2.402265072e-01f, 6.931471825e-01f, 1.000000000e+00f};
/* log2((1+s)/(1-s)) = s L(s^2) */
static const float LC[6] = {2.828850150e-01f, 3.199068010e-01f, 4.122092128e-01f, 5.770779252e-01f,
                            9.617967010e-01f, 2.885390043e+00f};

/* sin(pi a), a in [0, 1/2] */
static inline float sinpi_half(float a) {
    float x = a * a;
    float p = SPC[0];
    for (int k = 1; k < 5; ++k) p = fmaf(p, x, SPC[k]);
    return a * p;
}

float or_sinpi_sq(float v) {
    float r = v - rintf(v);
    float s = sinpi_half(fabsf(r));
    return s * s;
}

float or_sin2pi(float v) {
    float r = v - rintf(v);
    float a = fabsf(r);
    float b = fminf(a, 0.5f - a);
    float x = b * b;
    float p = SNC[0];
    for (int k = 1; k < 7; ++k) p = fmaf(p, x, SNC[k]);
    float s = b * p;
    return copysignf(s, r);
}

#define LOG2E_F 1.44269502e+00f
#define LN2_HI_F 6.93145752e-01f
#define LN2_LO_F 1.42860677e-06f

float or_expm1(float x) {
    if (x != x) return x;
    if (x < -17.5f) return -1.0f;
    if (x > 88.0f) return INFINITY;
    float n = rintf(x * LOG2E_F);
    float r = fmaf(-n, LN2_HI_F, x);
    r = fmaf(-n, LN2_LO_F, r);
    float p = EC[0];
    for (int k = 1; k < 7; ++k) p = fmaf(p, r, EC[k]);
    float r2 = r * r;
    float em1 = fmaf(r2, p, r);
    if (n == 0.0f) return em1;
    float t = pow2i((int)n);
    return fmaf(t, em1, t - 1.0f);
}

float or_exp2(float y) {
    if (y != y) return y;
    if (y > 128.0f) return INFINITY;
    if (y < -150.0f) return 0.0f;
    float n = rintf(y);
    float f = y - n;
    float p = X2C[0];
    for (int k = 1; k < 8; ++k) p = fmaf(p, f, X2C[k]);
    int ni = (int)n;
    int n1 = ni / 2;
    int n2 = ni - n1;
    return (p * pow2i(n1)) * pow2i(n2);
}

float or_log2(float x) {
    if (x != x || x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (isinf(x)) return x;
    int adj = 0;
    if (x < 1.17549435e-38f) { x = x * 16777216.0f; adj = -24; }
    uint32_t u = f2u(x);
    int e = (int)((u >> 23) & 0xFF) - 127 + adj;
    float m = u2f((u & 0x007FFFFFu) | 0x3F800000u);
    if (m > 1.41421354f) { m = m * 0.5f; e += 1; }
    float s = (m - 1.0f) / (m + 1.0f);
    float x2 = s * s;
    float p = LC[0];
    for (int k = 1; k < 6; ++k) p = fmaf(p, x2, LC[k]);
    float t = s * p;
    return (float)e + t;
}

float or_pow(float x, float y) {
    if (x == 0.0f) return (y > 0.0f) ? 0.0f : ((y == 0.0f) ? 1.0f : INFINITY);
    float l = or_log2(x);
    float ly = y * l;
    return or_exp2(ly);
}

/* Weierstrass per-dimension term, exact fractional parts of 3^k |z|. */
static const float P3F[21] = {1.f, 3.f, 9.f, 27.f, 81.f, 243.f, 729.f, 2187.f, 6561.f, 19683.f, 59049.f,
                              177147.f, 531441.f, 1594323.f, 4782969.f, 14348907.f, 43046721.f,
                              129140163.f, 387420489.f, 1162261467.f, 3486784401.f};

/* sin^2(pi t) = x Q_d(x), x = t^2, |t| <= 1/2: Q_d of degree d fitted to
 * sin^2(pi sqrt(x)) / x on [0, 1/4] (least squares reweighted toward the minimax relative
 * error; fit_coeffs.py): the engine's table (sg-emtf_amd/csrc/emtf_math.h kWQ5), d = 5. */
static const float WQ5[6] = {9.869604110717773f, -32.469661712646484f, 42.726924896240234f, -30.098546981811523f,
                             13.03709602355957f, -3.336198329925537f};
/* Fractional parts of 3^k |z| for 2^-32 <= |z| < 2^23 are carried as 0.64
 * fixed-point words y (frac = y / 2^64), so 3^(k+2) |z| mod 1 is y_{k+2} = 9 y_k mod 2^64
 * exactly; an even term k reads the top 32 bits as a signed word q, t = q 2^-32 in
 * [-1/2, 1/2) (q correctly rounded to binary32): t = frac or frac - 1, so
 * s = sin^2(pi t) = sin^2(pi frac) = x Q_5(x), with half the term's weight folded into
 * the coefficients (exact).  The odd term k + 1 is 2^-k s (3 - 4 s)^2 by the triple-angle
 * identity sin^2(3a) = sin^2(a) (3 - 4 sin^2(a))^2; terms are added in increasing k. */
float or_weier_term(float z) {
    if (z != z) return z;
    float a = fabsf(z);
    float acc = 0.0f;
    if (a >= 8388608.0f) return 0.0f;
    if (a < 2.32830644e-10f) { /* a < 2^-32: 3^k a < 1, no wrap */
        for (int k = 0; k <= 20; ++k) {
            float r = a * P3F[k];
            float s = sinpi_half(fminf(r, 1.0f - r));
            acc = fmaf(pow2i(1 - k), s * s, acc);
        }
        return acc;
    }
    uint32_t u = f2u(a);
    int ex = (int)((u >> 23) & 0xFF) - 127;    /* a in [2^ex, 2^(ex+1)), ex in [-32, 22] */
    uint64_t m = (uint64_t)((u & 0x007FFFFFu) | 0x00800000u);
    int s = 23 - ex;                            /* fraction bits, 1..55 */
    uint64_t y = m << (64 - s);                 /* integer part shifted out */
    for (int k = 0; k <= 20; k += 2) {
        int32_t q = (int32_t)(uint32_t)(y >> 32);
        float t = (float)q * 2.3283064365386963e-10f;
        float x = t * t;
        float hw = pow2i(-k); /* half the weight of term k */
        float c = WQ5[5] * hw;
        for (int i = 4; i >= 0; --i) c = fmaf(c, x, WQ5[i] * hw);
        float p = x * c;             /* 2^-k s, s = sin^2(pi 3^k z) */
        acc = fmaf(2.0f, p, acc);    /* term k: 2^(1-k) s */
        if (k == 20) break;
        float u = fmaf(p, -pow2i(k + 2), 3.0f); /* 3 - 4 s */
        float v = u * u;
        acc = fmaf(p, v, acc);       /* term k + 1: 2^-k s (3 - 4 s)^2 = 2^-k sin^2(pi 3^(k+1) z) */
        y = y + (y << 3);            /* 9 y mod 2^64 */
    }
    return acc;
}

/* ---- the reference's own formulas (evaluation(), SG-EMTF_DE.cu:834-1069) ---------- */

/* cos(pi y) = 1 - 2 sin^2(pi y / 2): y / 2 is exact and sin^2 reduces modulo 1 exactly,
 * so every finite y works; |y| >= 2^24 (even integers) gives exactly 1, as cos(pi y) does.
 * Stands in for CUDA's cospif (absolute error <= 2.4e-7). */
float or_cospi(float y) {
    float h = 0.5f * y;
    float s = or_sinpi_sq(h);
    float t = 2.0f * s;
    return 1.0f - t;
}

/* e^x (stands in for CUDA's expf; relative error <= 3e-7 on [-20, 2]) */
float or_exp(float x) { return or_expm1(x) + 1.0f; }

/* sum_{k=0}^{20} 0.5^k cospi(tv 3^k), tv = 2 (z + 0.5) rounded to binary32, the powers
 * 3^k from the binary32 recurrence p <- 3 p (equal to P3F), accumulated in increasing k
 * in binary32: the reference's temp_val / pow_val1 / pow_val2 loop, SG-EMTF_DE.cu:1012-1022.
 * The product tv 3^k is rounded to binary32 as the reference's is, which loses every
 * fraction bit once it passes 2^24 (cospi of an even integer: 1). */
float or_weier_ref_term(float z) {
    float zh = z + 0.5f;
    float tv = 2.0f * zh;
    float acc = 0.0f, p1 = 1.0f;
    for (int k = 0; k <= 20; ++k) {
        float y = tv * P3F[k];
        float c = or_cospi(y);
        float t = p1 * c;
        acc = acc + t;
        p1 = p1 * 0.5f;
    }
    return acc;
}

/* Weierstrass_para (SG-EMTF.cu:354-361): res += a cos(2 PI b 0.5) over k = 0..20 in
 * binary32 with PI = 3.1415926, the argument formed in double and rounded to binary32,
 * a = 0.5^k, b = 3^k (binary32 recurrence); the cosine correctly rounded to binary32
 * (the reference calls the host's cosf). */
float or_weier_ref_w(void) {
    float a = 1.0f, res = 0.0f;
    for (int k = 0; k <= 20; ++k) {
        double arg = 2.0 * 3.1415926 * (double)P3F[k] * 0.5;
        float c = (float)cos((double)(float)arg);
        float t = a * c;
        res = res + t;
        a = a * 0.5f;
    }
    return res;
}
