/*
 * Float32 elementary functions and the Philox4x32-10 counter RNG, written once
 * and compiled unchanged for the host (gcc/g++) and for gfx950 (hipcc), so the
 * CPU oracle and the HIP kernels produce bit-identical results.
 *
 * Build contract for bit-identical results on both sides:
 *   - no floating-point contraction (-ffp-contract=off), no fast-math;
 *   - IEEE float32 +, -, *, / and sqrtf (correctly rounded on x86-64 SSE and on
 *     gfx950 with hipcc's default correctly-rounded divide/sqrt);
 *   - every function below uses only those operations, floorf/fabsf and integer
 *     bit manipulation.
 *
 * Methods (all from the mathematics, no library code):
 *   sin/cos : Cody-Waite reduction by pi/2 (pi/2 split into 12-bit, 12-bit and
 *             24-bit pieces, so k*piece is exact for |k| < 2^12), then the
 *             Taylor series of sin to x^9 and of cos to x^10 on [-pi/4, pi/4]
 *             (truncation error < 3e-9 relative), quadrant by k mod 4.
 *   exp     : Cody-Waite reduction by ln 2 (same 3-piece split), Taylor series
 *             of e^r to r^8 on |r| <= ln2/2, scaling by 2^k through the
 *             exponent field in two factors (covers the subnormal range).
 *   log     : x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
 *             s = (m-1)/(m+1), odd series to s^13 (|s| <= 0.1716).
 *   tanh    : odd Taylor series (tangent numbers) to x^19 for |x| < 0.55,
 *             1 - 2/(e^{2|x|}+1) above, +-1 beyond 9.1.
 *   philox  : Philox4x32-10 of Salmon, Moraes, Dror and Shaw, "Parallel random
 *             numbers: as easy as 1, 2, 3" (SC'11): multipliers 0xD2511F53 and
 *             0xCD9E8D57, Weyl key increments 0x9E3779B9 and 0xBB67AE85.
 *   normal  : Box-Muller on two open-interval uniforms.
 */
#ifndef RACER_MATH_H
#define RACER_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RACER_FN __host__ __device__ static inline
#else
#include <math.h>
#include <string.h>
#define RACER_FN static inline
#endif

RACER_FN uint32_t racer_f2u(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    return u;
}

RACER_FN float racer_u2f(uint32_t u) {
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

/* Comparisons written out so the -0/NaN behaviour is the same everywhere. */
RACER_FN float racer_min(float a, float b) { return b < a ? b : a; }
RACER_FN float racer_max(float a, float b) { return a < b ? b : a; }
RACER_FN float racer_clamp(float x, float lo, float hi) { return racer_min(racer_max(x, lo), hi); }
RACER_FN float racer_abs(float x) { return racer_u2f(racer_f2u(x) & 0x7fffffffu); }

/* Round half away from zero to an integer-valued float (used for k in reductions). */
RACER_FN float racer_rint(float x) {
    return x >= 0.0f ? floorf(x + 0.5f) : -floorf(0.5f - x);
}

/* pi/2 = PIO2_1 + PIO2_2 + PIO2_3 (12 + 12 + 24 significant bits). */
#define RACER_PIO2_1 0x1.922p+0f
#define RACER_PIO2_2 (-0x1.2aep-18f)
#define RACER_PIO2_3 (-0x1.de973ep-31f)
#define RACER_2_OVER_PI 0x1.45f306p-1f
#define RACER_PI 0x1.921fb6p+1f
#define RACER_2PI 0x1.921fb6p+2f

/* sin and cos of |r| <= pi/4 (plus rounding slack). */
RACER_FN float racer_sin_kernel(float r) {
    float r2 = r * r;
    float p = (float)(1.0 / 362880.0);
    p = p * r2 + (float)(-1.0 / 5040.0);
    p = p * r2 + (float)(1.0 / 120.0);
    p = p * r2 + (float)(-1.0 / 6.0);
    return r + r * (r2 * p);
}

RACER_FN float racer_cos_kernel(float r) {
    float r2 = r * r;
    float p = (float)(-1.0 / 3628800.0);
    p = p * r2 + (float)(1.0 / 40320.0);
    p = p * r2 + (float)(-1.0 / 720.0);
    p = p * r2 + (float)(1.0 / 24.0);
    p = p * r2 + -0.5f;
    return 1.0f + r2 * p;
}

/* Reduce x to r in [-pi/4, pi/4] with x = k pi/2 + r; returns k mod 4. */
RACER_FN int racer_reduce_pio2(float x, float* r) {
    float k = racer_rint(x * RACER_2_OVER_PI);
    *r = ((x - k * RACER_PIO2_1) - k * RACER_PIO2_2) - k * RACER_PIO2_3;
    return ((int)k) & 3;
}

/* Both short kernels are evaluated and one is selected: on the GPU a choice
 * between calls compiles to a divergent branch per call, which costs a
 * serial per-car chain more than the other kernel's few instructions. */
RACER_FN float racer_sin(float x) {
    float r;
    int q = racer_reduce_pio2(x, &r);
    const float s = racer_sin_kernel(r), c = racer_cos_kernel(r);
    float v = (q & 1) ? c : s;
    return (q & 2) ? -v : v;
}

RACER_FN float racer_cos(float x) {
    float r;
    int q = racer_reduce_pio2(x, &r);
    const float s = racer_sin_kernel(r), c = racer_cos_kernel(r);
    float v = (q & 1) ? s : c;
    return ((q + 1) & 2) ? -v : v;
}

/* ln 2 = LN2_1 + LN2_2 + LN2_3 (12 + 12 + 24 significant bits). */
#define RACER_LN2_1 0x1.62ep-1f
#define RACER_LN2_2 0x1.0cp-15f
#define RACER_LN2_3 (-0x1.05c61p-29f)
#define RACER_INV_LN2 0x1.715476p+0f

/* e^x for finite x in [-103.972084, 88.7228394] (racer_exp's main path). */
RACER_FN float racer_exp_main(float x) {
    float k = racer_rint(x * RACER_INV_LN2);
    float r = ((x - k * RACER_LN2_1) - k * RACER_LN2_2) - k * RACER_LN2_3;
    float p = (float)(1.0 / 40320.0);
    p = p * r + (float)(1.0 / 5040.0);
    p = p * r + (float)(1.0 / 720.0);
    p = p * r + (float)(1.0 / 120.0);
    p = p * r + (float)(1.0 / 24.0);
    p = p * r + (float)(1.0 / 6.0);
    p = p * r + 0.5f;
    float e = 1.0f + (r + r * (r * p));
    int ki = (int)k;
    /* 2^k as two exact factors so k down to -150 stays representable. */
    int k1 = ki / 2;
    int k2 = ki - k1;
    return (e * racer_u2f((uint32_t)(k1 + 127) << 23)) * racer_u2f((uint32_t)(k2 + 127) << 23);
}

RACER_FN float racer_exp(float x) {
    if (!(x == x)) return x;
    if (x > 88.7228394f) return racer_u2f(0x7f800000u);
    if (x < -103.972084f) return 0.0f;
    return racer_exp_main(x);
}

RACER_FN float racer_log(float x) {
    if (!(x > 0.0f)) {
        if (x == 0.0f) return racer_u2f(0xff800000u);
        return racer_u2f(0x7fc00000u);
    }
    if (x == racer_u2f(0x7f800000u)) return x;
    uint32_t u = racer_f2u(x);
    int e = 0;
    if (u < 0x00800000u) { /* subnormal: scale by 2^24 first */
        u = racer_f2u(x * 0x1p24f);
        e = -24;
    }
    e += (int)(u >> 23) - 127;
    float m = racer_u2f((u & 0x007fffffu) | 0x3f800000u); /* [1, 2) */
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e += 1;
    }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = (float)(1.0 / 13.0);
    p = p * s2 + (float)(1.0 / 11.0);
    p = p * s2 + (float)(1.0 / 9.0);
    p = p * s2 + (float)(1.0 / 7.0);
    p = p * s2 + (float)(1.0 / 5.0);
    p = p * s2 + (float)(1.0 / 3.0);
    float logm = 2.0f * s + 2.0f * s * (s2 * p);
    float fe = (float)e;
    return (fe * RACER_LN2_3 + (fe * RACER_LN2_2 + logm)) + fe * RACER_LN2_1;
}

RACER_FN float racer_tanh(float x) {
    float a = racer_abs(x);
    float y;
    if (a < 0.55f) {
        float x2 = a * a;
        float p = (float)(-443861162.0 / 1856156927625.0);
        p = p * x2 + (float)(6404582.0 / 10854718875.0);
        p = p * x2 + (float)(-929569.0 / 638512875.0);
        p = p * x2 + (float)(21844.0 / 6081075.0);
        p = p * x2 + (float)(-1382.0 / 155925.0);
        p = p * x2 + (float)(62.0 / 2835.0);
        p = p * x2 + (float)(-17.0 / 315.0);
        p = p * x2 + (float)(2.0 / 15.0);
        p = p * x2 + (float)(-1.0 / 3.0);
        y = a + a * (x2 * p);
    } else if (a < 9.1f) {
        y = 1.0f - 2.0f / (racer_exp(2.0f * a) + 1.0f);
    } else {
        y = (a == a) ? 1.0f : a;
    }
    return x < 0.0f ? -y : y;
}

/* racer_tanh with no branches: both finite-range paths (racer_tanh's own
 * operations) are evaluated and one is selected, so the result is racer_tanh's
 * bit for bit, and many independent evaluations can be interleaved by the
 * compiler (where racer_tanh's branches serialise them). The exp path's
 * argument is clamped to 18.2 (NaN included), which changes nothing for the
 * |x| < 9.1 it serves and keeps racer_exp_main in its range otherwise. */
RACER_FN float racer_tanh_select(float x) {
    const float a = racer_abs(x);
    const float x2 = a * a;
    float p = (float)(-443861162.0 / 1856156927625.0);
    p = p * x2 + (float)(6404582.0 / 10854718875.0);
    p = p * x2 + (float)(-929569.0 / 638512875.0);
    p = p * x2 + (float)(21844.0 / 6081075.0);
    p = p * x2 + (float)(-1382.0 / 155925.0);
    p = p * x2 + (float)(62.0 / 2835.0);
    p = p * x2 + (float)(-17.0 / 315.0);
    p = p * x2 + (float)(2.0 / 15.0);
    p = p * x2 + (float)(-1.0 / 3.0);
    const float small = a + a * (x2 * p);
    float e2 = 2.0f * a;
    e2 = e2 < 18.2f ? e2 : 18.2f;
    const float middle = 1.0f - 2.0f / (racer_exp_main(e2) + 1.0f);
    /* one select per statement: a conditional expression nested in another is
     * emitted as control flow, and the paths' arithmetic then sinks into it */
    const float big = (a == a) ? 1.0f : a;
    const float upper = a < 9.1f ? middle : big;
    const float y = a < 0.55f ? small : upper;
    return x < 0.0f ? -y : y;
}

/* A branch-free tanh for the bf16 surrogate's hidden layers, whose outputs
 * are rounded to bf16 (8 bits) or feed a learned output layer: the odd series
 * of racer_tanh with fused multiply-adds below 0.55, 1 - 2 / (2^(2 a log2 e)
 * + 1) above it with the GPU's approximate exp2 and reciprocal (libm exp2f and
 * a division on the host). About a quarter of racer_tanh's instructions; a
 * few ulp from tanh, not bitwise racer_tanh. */
#if defined(__HIP_DEVICE_COMPILE__)
#define RACER_EXP2_APPROX(v) __builtin_amdgcn_exp2f(v)
#define RACER_RCP_APPROX(v) __builtin_amdgcn_rcpf(v)
#else
#define RACER_EXP2_APPROX(v) exp2f(v)
#define RACER_RCP_APPROX(v) (1.0f / (v))
#endif
RACER_FN float racer_tanh_mlp(float x) {
    const float a = racer_abs(x);
    const float x2 = a * a;
    float p = (float)(-443861162.0 / 1856156927625.0);
    p = fmaf(p, x2, (float)(6404582.0 / 10854718875.0));
    p = fmaf(p, x2, (float)(-929569.0 / 638512875.0));
    p = fmaf(p, x2, (float)(21844.0 / 6081075.0));
    p = fmaf(p, x2, (float)(-1382.0 / 155925.0));
    p = fmaf(p, x2, (float)(62.0 / 2835.0));
    p = fmaf(p, x2, (float)(-17.0 / 315.0));
    p = fmaf(p, x2, (float)(2.0 / 15.0));
    p = fmaf(p, x2, (float)(-1.0 / 3.0));
    const float small = fmaf(a, x2 * p, a);
    const float t = RACER_EXP2_APPROX(a * 2.88539008f);  /* e^(2a) */
    const float large = fmaf(-2.0f, RACER_RCP_APPROX(t + 1.0f), 1.0f);
    const float y = a < 0.55f ? small : large;
    return x < 0.0f ? -y : y;
}

/* floorf(a / b + 0.5f) with a / b correctly rounded, as a reciprocal multiply
 * in the common case. inv_b = RN(1/b). q0 = RN(a * inv_b) differs from
 * q = RN(a / b) by at most 3 ulp(q) (relative errors of 2^-24 from inv_b, 2^-24
 * from the product, 2^-24 from q itself), and s0 = RN(q0 + 0.5) from
 * RN(q + 0.5) by at most 3 ulp(q) + ulp(q + 0.5) <= 5 ulp(max(|q0|, 1)). So
 * when s0 is more than (|q0| + 1) 2^-20 - 2^-25 (> 7.75 such ulps) away from
 * both neighbouring integers, both floors agree; otherwise the true division
 * runs. The test: s0 - f0 is exact, lim = RN(0.5 - (|q0| + 1) 2^-20) is one
 * fused rounding (error <= 2^-26 while lim > 0), |(s0 - f0) - 0.5| is off by
 * at most 2^-26, so passing it puts s0 - f0 strictly inside
 * ((|q0| + 1) 2^-20 - 2^-25, 1 - that). A NaN fails it; |q0| >= 2^19 - 1
 * (lim <= 0) always takes the division. */
RACER_FN float racer_nearest_index(float a, float b, float inv_b) {
    float q0 = a * inv_b;
    float s0 = q0 + 0.5f;
    float f0 = floorf(s0);
    float lim = fmaf(racer_abs(q0), -0x1p-20f, 0.5f - 0x1p-20f);
    if (racer_abs((s0 - f0) - 0.5f) < lim) return f0;
    return floorf(a / b + 0.5f);
}

/* The same with the limit supplied by the caller for a run of calls:
 * racer_nearest_limit(Q) = RN(0.5 - (Q + 1) 2^-20) is at most the per-call
 * limit of every q0 with |q0| <= Q (rounding is monotone), so for such calls
 * the test is stricter and the result the same. */
RACER_FN float racer_nearest_limit(float Q) { return fmaf(Q, -0x1p-20f, 0.5f - 0x1p-20f); }
RACER_FN float racer_nearest_index_lim(float a, float b, float inv_b, float lim) {
    float q0 = a * inv_b;
    float s0 = q0 + 0.5f;
    float f0 = floorf(s0);
    if (racer_abs((s0 - f0) - 0.5f) < lim) return f0;
    return floorf(a / b + 0.5f);
}

/* a / b correctly rounded (to nearest even) from y = RN(1 / b), with one
 * multiply and four FMAs, for b in [2^-60, 2^60] and a = +0 or |a| in
 * [2^-60, 2^60] (no step then overflows, and every residual is a normal
 * number). This is Markstein's FMA division (P. Markstein, "Computation of
 * elementary functions on the IBM RISC System/6000 processor", IBM J. Res.
 * Dev. 34(1), 1990; J.-M. Muller et al., Handbook of Floating-Point
 * Arithmetic, FMA-based division):
 *  - q0 = RN(a y) is within 2^-23 (1 + 2^-24) relative of a / b (y's and the
 *    product's roundings), so within two ulps of it;
 *  - q0 + RN(a - b q0) y differs from a / b by (a / b - q0)(e1 + e3 + e1 e3),
 *    |e1|, |e3| <= 2^-24: at most 2^-45.9 relative, so its rounding q1 is a
 *    faithful rounding of a / b (one of the two floats around it);
 *  - with q1 faithful and y within half an ulp of 1 / b, r = a - b q1 is
 *    exactly representable and RN(q1 + r y) = RN(a / b) (Markstein's
 *    theorem).
 * a = -0 would give +0 where the division gives -0. */
RACER_FN float racer_div_by_reciprocal(float a, float b, float y) {
    const float q0 = a * y;
    const float q1 = fmaf(fmaf(-q0, b, a), y, q0);
    return fmaf(fmaf(-q1, b, a), y, q1);
}

/* Map.at's nearest cell floorf(a / b + 0.5f) from y = RN(1 / b) with one
 * correction step: q1 = RN(q0 + RN(a - b q0) y) is a faithful rounding of a /
 * b (the first two steps above; b in [2^-20, 2^20], |a| <= 2^60), so it is
 * RN(a / b) or the float on the other side of a / b. g(x) = floorf(x + 0.5f)
 * is non-decreasing, so g(q1) and g(RN(a / b)) can differ only where a jump
 * of g lies between those two neighbours, that is for a / b within one ulp of
 * a jump point; every jump point is within two ulps of a half-integer N - 0.5.
 * racer_nearest_form_bad checks every float a whose quotient lies near N -
 * 0.5 for a range of N; where it finds none that differs, this form is the
 * division's index for every a with a / b inside that range of half-integers
 * (and outside it on the same side of the range, since q1 is faithful). */
RACER_FN float racer_nearest_by_reciprocal1(float a, float b, float y) {
    const float q0 = a * y;
    return floorf(fmaf(fmaf(-q0, b, a), y, q0) + 0.5f);
}

/* The count of floats a with (N - 0.5) b - 256 ulp <= a <= (N - 0.5) b + 256
 * ulp, for N in [nlo, nhi], whose index by racer_nearest_by_reciprocal1 (form
 * 1) or by RN(a y) alone (form 2, kept to show that the check finds the
 * failures of a form that has them) differs from floorf(a / b + 0.5f). The
 * window is far wider than the two ulps of quotient that matter: a / b moves
 * by at most two ulps of the quotient per ulp of a. */
RACER_FN long racer_nearest_form_bad(float b, int nlo, int nhi, int form) {
    const float y = 1.0f / b;
    long bad = 0;
    for (int N = nlo; N <= nhi; ++N) {
        const float c = (float)(((double)N - 0.5) * (double)b);
        /* step through the floats from 256 below c to 256 above it */
        for (int k = -256; k <= 256; ++k) {
            const uint32_t u = racer_f2u(c);
            const int64_t o = (u & 0x80000000u) ? -(int64_t)(u & 0x7fffffffu) : (int64_t)u;
            const int64_t ok = o + k;
            const float a = racer_u2f(ok < 0 ? (uint32_t)(-ok) | 0x80000000u : (uint32_t)ok);
            const float want = floorf(a / b + 0.5f);
            const float got = form == 1 ? racer_nearest_by_reciprocal1(a, b, y) : floorf(a * y + 0.5f);
            if (racer_f2u(want) != racer_f2u(got)) ++bad;
        }
    }
    return bad;
}

#if defined(__HIPCC__)
/* Branch-free global loads through gfx9 buffer descriptors: an offset at or
 * beyond `bytes` reads as zero in hardware, so guarded loads need no branch
 * (a branch around a load makes the compiler wait for every load in flight). */
#define RACER_OOB 0xFFFFFFF0u
__device__ static inline __amdgpu_buffer_rsrc_t racer_rsrc(const void* base, uint64_t bytes) {
    uint32_t n = bytes >= 0xFFFFFFF0ull ? 0xFFFFFFF0u : (uint32_t)bytes;
    return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)n, 0x00020000);
}
__device__ static inline float racer_ld_f32(__amdgpu_buffer_rsrc_t r, uint32_t byte_offset) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)byte_offset, 0, 0));
}
__device__ static inline uint4 racer_ld_b128(__amdgpu_buffer_rsrc_t r, uint32_t byte_offset) {
    return __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)byte_offset, 0, 0));
}
__device__ static inline float2 racer_ld_f32x2(__amdgpu_buffer_rsrc_t r, uint32_t byte_offset) {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)byte_offset, 0, 0));
}
__device__ static inline uint16_t racer_ld_u16(__amdgpu_buffer_rsrc_t r, uint32_t byte_offset) {
    return (uint16_t)__builtin_amdgcn_raw_buffer_load_b16(r, (int)byte_offset, 0, 0);
}
#endif

/* ---- Philox4x32-10 ---- */
typedef struct {
    uint32_t v[4];
} racer_u32x4;

RACER_FN racer_u32x4 racer_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    racer_u32x4 out = {{c0, c1, c2, c3}};
    return out;
}

/* One block of the stream (seed, subsequence, offset): key = seed, counter =
 * (offset lo, offset hi, subsequence lo, subsequence hi). */
RACER_FN racer_u32x4 racer_random(uint64_t seed, uint64_t subsequence, uint64_t offset) {
    return racer_philox((uint32_t)offset, (uint32_t)(offset >> 32), (uint32_t)subsequence,
                        (uint32_t)(subsequence >> 32), (uint32_t)seed, (uint32_t)(seed >> 32));
}

/* Uniform in [0, 1) from the top 24 bits. */
RACER_FN float racer_uniform(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }

/* Uniform in (0, 1), never 0 or 1. */
RACER_FN float racer_uniform_open(uint32_t x) { return ((float)(x >> 8) + 0.5f) * 0x1p-24f; }

/* Four standard normals from one Philox block by Box-Muller. */
RACER_FN void racer_normal4(racer_u32x4 r, float out[4]) {
    for (int h = 0; h < 2; ++h) {
        float u = racer_uniform_open(r.v[2 * h]);
        float w = racer_uniform_open(r.v[2 * h + 1]);
        float radius = sqrtf(-2.0f * racer_log(u));
        float angle = RACER_2PI * w;
        out[2 * h] = radius * racer_cos(angle);
        out[2 * h + 1] = radius * racer_sin(angle);
    }
}

#endif
