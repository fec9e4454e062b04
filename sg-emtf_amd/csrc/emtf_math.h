// emtf_math.h -- device arithmetic of the island engine (gfx950).
//
// Every function here is specified bit-for-bit: IEEE binary32 with
// round-to-nearest-even, explicit fmaf only (the library is compiled with
// -ffp-contract=off), correctly rounded division and sqrt.  The same
// definitions evaluated on any IEEE host reproduce the device results exactly,
// which is what the parity tests check.
//
// Random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11) with
//   key     = { seed lo32, seed hi32 }
//   counter = { block, purpose<<24 | sub<<16 | index, generation, task }.
// It replaces the per-thread cuRAND XORWOW states of the reference
// (proposed model/SG-EMTF.cu:163-179); draws are addressed by what they are
// for, never by which thread consumes them.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace emtf {

enum Purpose : uint32_t {
    P_INIT = 1, P_DE_IDX = 2, P_DE_CR = 3, P_GA_MATE = 4, P_GA_SBX = 5,
    P_GA_PM = 6, P_MIG = 7, P_XFER = 8, P_ISLX = 9
};

struct U4 { uint32_t x, y, z, w; };

__device__ __forceinline__ U4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                     uint32_t k0, uint32_t k1) {
    // the key schedule is re-derived per block with two scalar adds a round; without
    // this the compiler keeps all twenty round keys live across the generation loop,
    // where they spill and come back as VALU lane reads
    asm volatile("" : "+s"(k0), "+s"(k1));
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return U4{c0, c1, c2, c3};
}

struct Rng {
    uint32_t k0, k1, task;
    __device__ __forceinline__ U4 block(uint32_t gen, uint32_t purpose, uint32_t index, uint32_t sub,
                                        uint32_t blk) const {
        return philox(blk, (purpose << 24) | ((sub & 0xFFu) << 16) | (index & 0xFFFFu), gen, task, k0, k1);
    }
    __device__ __forceinline__ uint32_t word(uint32_t gen, uint32_t purpose, uint32_t index, uint32_t sub,
                                             uint32_t w) const {
        const U4 b = block(gen, purpose, index, sub, w >> 2);
        const uint32_t s = w & 3u;
        return s == 0 ? b.x : (s == 1 ? b.y : (s == 2 ? b.z : b.w));
    }
};

__device__ __forceinline__ uint32_t pick(const U4& b, uint32_t s) {
    return s == 0 ? b.x : (s == 1 ? b.y : (s == 2 ? b.z : b.w));
}

// (0,1]: ((w >> 8) + 1) * 2^-24, exact in binary32 (reference: curand_uniform's range)
__device__ __forceinline__ float uniform01(uint32_t w) {
    return (float)((w >> 8) + 1u) * 5.9604644775390625e-08f;
}
// (0,1) symmetric: (2 (w >> 9) + 1) * 2^-24
__device__ __forceinline__ float uniform_open(uint32_t w) {
    return (float)(2u * (w >> 9) + 1u) * 5.9604644775390625e-08f;
}

__device__ __forceinline__ float pow2i(int n) { return __int_as_float((n + 127) << 23); }

// ---- polynomials (Chebyshev fits rounded to binary32) -----------------------
// sin(pi a) = a SP(a^2), a in [0, 1/2]; max rel. error 3.6e-8
__device__ __forceinline__ float sinpi_half(float a) {
    const float x = a * a;
    float p = 7.765591145e-02f;
    p = fmaf(p, x, -5.982903838e-01f);
    p = fmaf(p, x, 2.550077438e+00f);
    p = fmaf(p, x, -5.167710304e+00f);
    p = fmaf(p, x, 3.141592741e+00f);
    return a * p;
}
// sin(2 pi b) = b SN(x), x = b^2, b in [0, 1/4]; max rel. error 5.3e-8
__device__ __forceinline__ float poly_sn(float x) {
    float p = 3.665855169e+00f;
    p = fmaf(p, x, -1.508148289e+01f);
    p = fmaf(p, x, 4.205813599e+01f);
    p = fmaf(p, x, -7.670584869e+01f);
    p = fmaf(p, x, 8.160524750e+01f);
    p = fmaf(p, x, -4.134170151e+01f);
    p = fmaf(p, x, 6.283185482e+00f);
    return p;
}

// sin^2(pi v)
__device__ __forceinline__ float sinpi_sq(float v) {
    const float r = v - __builtin_rintf(v);
    const float sp = sinpi_half(__builtin_fabsf(r));
    return sp * sp;
}
// sin(2 pi v)
__device__ __forceinline__ float sin2pi(float v) {
    const float r = v - __builtin_rintf(v);
    const float a = __builtin_fabsf(r);
    const float b = __builtin_fminf(a, 0.5f - a);
    const float x = b * b;
    const float s = b * poly_sn(x);
    return __builtin_copysignf(s, r);
}

// e^x - 1: Cody-Waite reduction by ln2 (hi/lo split), polynomial on |r| <= ln2/2
__device__ __forceinline__ float expm1_f(float x) {
    if (x != x) return x;
    if (x < -17.5f) return -1.0f;
    if (x > 88.0f) return __builtin_inff();
    const float n = __builtin_rintf(x * 1.44269502e+00f);
    float r = fmaf(-n, 6.93145752e-01f, x);
    r = fmaf(-n, 1.42860677e-06f, r);
    float p = 2.485957884e-05f;
    p = fmaf(p, r, 1.989927405e-04f);
    p = fmaf(p, r, 1.388885430e-03f);
    p = fmaf(p, r, 8.333298378e-03f);
    p = fmaf(p, r, 4.166666791e-02f);
    p = fmaf(p, r, 1.666666716e-01f);
    p = fmaf(p, r, 5.000000000e-01f);
    const float r2 = r * r;
    const float em1 = fmaf(r2, p, r);
    if (n == 0.0f) return em1;
    const float t = pow2i((int)n);
    return fmaf(t, em1, t - 1.0f);
}

__device__ __forceinline__ float exp2_f(float y) {
    if (y != y) return y;
    if (y > 128.0f) return __builtin_inff();
    if (y < -150.0f) return 0.0f;
    const float n = __builtin_rintf(y);
    const float f = y - n;
    float p = 1.530369991e-05f;
    p = fmaf(p, f, 1.546972926e-04f);
    p = fmaf(p, f, 1.333347871e-03f);
    p = fmaf(p, f, 9.618025273e-03f);
    p = fmaf(p, f, 5.550410971e-02f);
    p = fmaf(p, f, 2.402265072e-01f);
    p = fmaf(p, f, 6.931471825e-01f);
    p = fmaf(p, f, 1.000000000e+00f);
    const int ni = (int)n;
    const int n1 = ni / 2;
    const int n2 = ni - n1;
    return (p * pow2i(n1)) * pow2i(n2);
}

__device__ __forceinline__ float log2_f(float x) {
    if (x != x || x < 0.0f) return __builtin_nanf("");
    if (x == 0.0f) return -__builtin_inff();
    if (__builtin_isinf(x)) return x;
    int adj = 0;
    if (x < 1.17549435e-38f) { x = x * 16777216.0f; adj = -24; }
    const uint32_t u = __float_as_uint(x);
    int e = (int)((u >> 23) & 0xFFu) - 127 + adj;
    float m = __uint_as_float((u & 0x007FFFFFu) | 0x3F800000u);
    if (m > 1.41421354f) { m = m * 0.5f; e += 1; }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float x2 = s * s;
    float p = 2.828850150e-01f;
    p = fmaf(p, x2, 3.199068010e-01f);
    p = fmaf(p, x2, 4.122092128e-01f);
    p = fmaf(p, x2, 5.770779252e-01f);
    p = fmaf(p, x2, 9.617967010e-01f);
    p = fmaf(p, x2, 2.885390043e+00f);
    const float t = s * p;
    return (float)e + t;
}

// x^y for x >= 0 as 2^(y log2 x)
__device__ __forceinline__ float pow_f(float x, float y) {
    if (x == 0.0f) return (y > 0.0f) ? 0.0f : ((y == 0.0f) ? 1.0f : __builtin_inff());
    const float l = log2_f(x);
    const float ly = y * l;
    return exp2_f(ly);
}

// sin^2(pi t) = x Q_d(x), x = t^2, t in [-1/2, 1/2]: Q_d of degree d fitted to
// sin^2(pi sqrt(x)) / x on [0, 1/4] by least squares reweighted toward the minimax of the
// relative error (Lawson's iteration; oracle/fit_coeffs.py); largest relative error of the
// binary32 Horner form: d = 5: 1.9e-7 (also at x -> 0, so a sum of terms keeps its
// relative accuracy near the optimum).
constexpr float kWQ5[6] = {9.869604110717773f, -32.469661712646484f, 42.726924896240234f, -30.098546981811523f,
                           13.03709602355957f, -3.336198329925537f};

// Weierstrass term sum_{k=0}^{20} 2^(1-k) sin^2(pi 3^k z).  For 2^-32 <= |z| < 2^23
// the fractional part of 3^k |z| is carried exactly as a 0.64 fixed-point word (9x mod 1,
// two terms on, is one 64-bit shift-add).  An even term k reads the word's top 32 bits as
// a signed word: t = q 2^-32 in [-1/2, 1/2) differs from the fraction by 0 or 1, so
// s = sin^2(pi t) = x Q_5(x), the weight folded into the coefficients (exact: powers of
// two).  The odd term k + 1 follows by the triple-angle identity sin^2(3a) = s (3 - 4 s)^2
// in three operations: its absolute error is at most 9 times that of s and it weighs
// half as much, so each pair of terms keeps its relative error within about 4 times the
// polynomial's.  Against 40-digit sums: 6.1e-7 relative at most over 4000 points (1.1e-6
// with every term from a degree-4 polynomial of its own).
//
// The fixed-point path of weier_term for a = |z| in [2^-32, 2^23) (ok: a is), without
// branches; +0 where ok is false, which is the term for a = 0 and a >= 2^23 (not for NaN
// or 0 < a < 2^-32, which weier_term sends to its small-|z| branch).
__device__ __forceinline__ float weier_core(float a, bool ok) {
    const uint32_t u = __float_as_uint(a);
    const int ex = (int)((u >> 23) & 0xFFu) - 127;
    const uint64_t m = (uint64_t)((u & 0x007FFFFFu) | 0x00800000u);
    uint64_t y = m << ((41 + ex) & 63);  // 64 - (23 - ex) fraction bits: 9 .. 63 where ok
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k <= 20; k += 2) {
        // the conversion of the high word as one 32-bit instruction (left to itself the
        // backend widens it to a 64-bit conversion)
        float qf;
        asm("v_cvt_f32_i32 %0, %1" : "=v"(qf) : "v"((uint32_t)(y >> 32)));
        const float t = qf * 2.3283064365386963e-10f;
        const float x = t * t;
        const float hw = pow2i(-k);  // half the weight of term k
        float c = kWQ5[5] * hw;
#pragma unroll
        for (int i = 4; i >= 0; --i) c = fmaf(c, x, kWQ5[i] * hw);
        const float p = x * c;          // 2^-k s, s = sin^2(pi 3^k z)
        acc = fmaf(2.0f, p, acc);       // term k: 2^(1-k) s
        if (k == 20) break;
        const float u = fmaf(p, -pow2i(k + 2), 3.0f);  // 3 - 4 s
        const float v = u * u;
        acc = fmaf(p, v, acc);  // term k + 1: 2^-k s (3 - 4 s)^2 = 2^-k sin^2(pi 3^(k+1) z)
        // y <- 9 y mod 2^64 in one VALU op (the backend otherwise emits 32-bit multiplies
        // and register moves per step)
        asm("v_lshl_add_u64 %0, %1, 3, %1" : "=v"(y) : "v"(y));
    }
    return ok ? acc : 0.0f;
}

__device__ __forceinline__ float weier_term(float z) {
    if (z != z) return z;
    const float a = __builtin_fabsf(z);
    float acc = 0.0f;
    if (a >= 8388608.0f) return 0.0f;
    if (a < 2.32830644e-10f) {
        // 3^k as binary32 (exact for k <= 15, correctly rounded above); 3^k a < 1
        constexpr float P3F[21] = {1.f, 3.f, 9.f, 27.f, 81.f, 243.f, 729.f, 2187.f, 6561.f, 19683.f,
                                   59049.f, 177147.f, 531441.f, 1594323.f, 4782969.f, 14348907.f,
                                   43046721.f, 129140163.f, 387420489.f, 1162261467.f, 3486784401.f};
#pragma unroll
        for (int k = 0; k <= 20; ++k) {
            const float r = a * P3F[k];
            const float sp = sinpi_half(__builtin_fminf(r, 1.0f - r));
            acc = fmaf(pow2i(1 - k), sp * sp, acc);
        }
        return acc;
    }
    return weier_core(a, true);
}

// ---- the reference's own formulas (evaluation(), SG-EMTF_DE.cu:834-1069) ----------
// cos(pi y) = 1 - 2 sin^2(pi y / 2) for any finite binary32 y (exact reduction; |y| >= 2^24
// gives 1); stands in for CUDA's cospif
__device__ __forceinline__ float cospi_f(float y) {
    const float h = 0.5f * y;
    const float s = sinpi_sq(h);
    const float t = 2.0f * s;
    return 1.0f - t;
}

// e^x = expm1(x) + 1; stands in for CUDA's expf
__device__ __forceinline__ float exp_f(float x) { return expm1_f(x) + 1.0f; }

// sum_{k=0}^{20} 0.5^k cospi(tv 3^k), tv = 2 (z + 0.5) and each product rounded to binary32
// as the reference's cospif argument is (SG-EMTF_DE.cu:1012-1022); 3^k by the binary32
// recurrence p <- 3 p, whose values are the correctly rounded powers
__device__ __forceinline__ float weier_ref_term(float z) {
    constexpr float P3R[21] = {1.f, 3.f, 9.f, 27.f, 81.f, 243.f, 729.f, 2187.f, 6561.f, 19683.f,
                               59049.f, 177147.f, 531441.f, 1594323.f, 4782969.f, 14348907.f,
                               43046721.f, 129140163.f, 387420489.f, 1162261467.f, 3486784401.f};
    const float zh = z + 0.5f;
    const float tv = 2.0f * zh;
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k <= 20; ++k) {
        const float y = tv * P3R[k];
        const float c = cospi_f(y);
        const float t = pow2i(-k) * c;
        acc = acc + t;
    }
    return acc;
}

// 16-bit uniform u = (v + 1) 2^-16 in (0,1]: half 0 = high 16 bits of w, half 1 = low.
// u > CR  <=>  v >= floor(CR 2^16) for every binary32 CR in [0,1] (both sides exact), so
// the crossover test compares the raw 16-bit value with a precomputed threshold.
__device__ __forceinline__ uint32_t bits16(uint32_t w, int half) { return half ? (w & 0xFFFFu) : (w >> 16); }

// n mod d for a divisor fixed per task: Granlund-Montgomery round-up multiplier,
// exact for every 32-bit n (d >= 1).
struct FastMod {
    uint32_t d, mul;
    int sh;  // -1: d == 1
    __device__ __forceinline__ void init(uint32_t dv) {
        d = dv;
        if (dv <= 1) { mul = 0; sh = -1; return; }
        const int l = 32 - __builtin_clz(dv - 1);                     // ceil(log2 d)
        mul = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << l) - dv)) / dv + 1);
        sh = l;
    }
    __device__ __forceinline__ uint32_t mod(uint32_t n) const {
        if (sh < 0) return 0u;
        const uint32_t t1 = __umulhi(n, mul);
        const uint32_t q = (t1 + ((n - t1) >> 1)) >> (sh - 1);
        return n - q * d;
    }
};

}  // namespace emtf
