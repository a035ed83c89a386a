// gs_math.h -- float32 transcendental functions shared by the gfx950 kernels
// and the host code that checks them.
//
// Every function here is built only from IEEE-754 operations that are
// correctly rounded on both gfx950 and x86-64 (+, -, *, /, sqrt, fma,
// round-to-nearest-even) plus integer bit manipulation.  Compiled with
// -ffp-contract=off on both sides, the same input therefore gives the same
// output bits on the GPU and on the CPU.  That is what lets the per-pixel
// blend decisions (alpha < 1/255, T < 1e-4) of the GPU path agree exactly with
// a CPU restatement of the rasterizer.
//
// Methods (textbook; J.-M. Muller, "Elementary Functions", chapters 3 and 11):
//   gs_expf   Cody-Waite reduction x = n*ln2 + r, e^r = 1 + r + r^2 q(r)
//   gs_logf   x = 2^e * m, m in [sqrt(1/2), sqrt(2)), log1p(f) = f + f^2 Q(f)
//   gs_atan2f octant reduction to a in [0,1], atan(a) = a + a^3 P(a^2)
//   gs_asinf  asin(x) = x + x^3 S(x^2) on [0, 1/2];
//             asin(x) = pi/2 - 2 asin(sqrt((1 - x) / 2)) above
// The polynomial coefficients come from gen_gs_math_coeffs.py (near-minimax
// relative-error fits); accuracy is measured against double precision by the
// test suite (all within 2 ulp over the tested ranges).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define GS_HD __host__ __device__ inline
#else
#define GS_HD static inline
#endif

// ---- polynomial coefficients (generated by gen_gs_math_coeffs.py) ----
#define GS_EXP_Q_0 0x1.0000000000000p-1f  /* 0.499999994195 */
#define GS_EXP_Q_1 0x1.5554de0000000p-3f  /* 0.166665774514 */
#define GS_EXP_Q_2 0x1.5555ea0000000p-5f  /* 0.0416669426924 */
#define GS_EXP_Q_3 0x1.120afc0000000p-7f  /* 0.00836312576496 */
#define GS_EXP_Q_4 0x1.6c37ec0000000p-10f  /* 0.00138938315321 */
#define GS_LOG_Q_0 -0x1.fffffe0000000p-2f  /* -0.499999976593 */
#define GS_LOG_Q_1 0x1.5555560000000p-2f  /* 0.333333339083 */
#define GS_LOG_Q_2 -0x1.00020e0000000p-2f  /* -0.250007823186 */
#define GS_LOG_Q_3 0x1.999e520000000p-3f  /* 0.20000900083 */
#define GS_LOG_Q_4 -0x1.5481b80000000p-3f  /* -0.166263049023 */
#define GS_LOG_Q_5 0x1.2310760000000p-3f  /* 0.142121237295 */
#define GS_LOG_Q_6 -0x1.0d13ba0000000p-3f  /* -0.131385281011 */
#define GS_LOG_Q_7 0x1.039fb60000000p-3f  /* 0.126769462151 */
#define GS_LOG_Q_8 -0x1.36c5780000000p-4f  /* -0.075871916613 */
#define GS_ATAN_P_0 -0x1.5555540000000p-2f  /* -0.333333316425 */
#define GS_ATAN_P_1 0x1.99982e0000000p-3f  /* 0.199997284768 */
#define GS_ATAN_P_2 -0x1.246bb00000000p-3f  /* -0.142783516781 */
#define GS_ATAN_P_3 0x1.c3dfde0000000p-4f  /* 0.110320917165 */
#define GS_ATAN_P_4 -0x1.62531e0000000p-4f  /* -0.0865050477669 */
#define GS_ATAN_P_5 0x1.feed3c0000000p-5f  /* 0.062368983369 */
#define GS_ATAN_P_6 -0x1.2499d40000000p-5f  /* -0.0357178821944 */
#define GS_ATAN_P_7 0x1.b793dc0000000p-7f  /* 0.0134148426016 */
#define GS_ATAN_P_8 -0x1.35de580000000p-9f  /* -0.00236410926598 */
#define GS_ASIN_S_0 0x1.5555540000000p-3f  /* 0.166666663386 */
#define GS_ASIN_S_1 0x1.3334300000000p-4f  /* 0.0750009431412 */
#define GS_ASIN_S_2 0x1.6d5be00000000p-5f  /* 0.0445994749825 */
#define GS_ASIN_S_3 0x1.fd8a2e0000000p-6f  /* 0.0310998398876 */
#define GS_ASIN_S_4 0x1.1908e60000000p-6f  /* 0.0171529998826 */
#define GS_ASIN_S_5 0x1.13f2360000000p-5f  /* 0.0336848294275 */

// ln2 = GS_LN2_HI + GS_LN2_LO (hi is float(ln2))
#define GS_LN2_HI 0x1.62e430p-1f
#define GS_LN2_LO -0x1.05c610p-29f
#define GS_LOG2E 0x1.715476p+0f
// pi/2 and pi split the same way
#define GS_PIO2_HI 0x1.921fb6p+0f
#define GS_PIO2_LO -0x1.777a5cp-25f
#define GS_PI_HI 0x1.921fb6p+1f
#define GS_PI_LO -0x1.777a5cp-24f

typedef union { float f; uint32_t u; } gs_f32_bits;
GS_HD float gs_as_float(uint32_t u) { gs_f32_bits b; b.u = u; return b.f; }
GS_HD uint32_t gs_as_uint(float f) { gs_f32_bits b; b.f = f; return b.u; }

// 2^k for k in [-126, 127], built from the exponent field.
GS_HD float gs_pow2i(int k) { return gs_as_float((uint32_t)(k + 127) << 23); }

GS_HD float gs_expf(float x)
{
    if (x != x) return x;
    if (x >= 88.72283935546875f) return __builtin_huge_valf();
    if (x < -104.0f) return 0.0f;
    const float n = __builtin_rintf(x * GS_LOG2E);
    float r = __builtin_fmaf(-n, GS_LN2_HI, x);
    r = __builtin_fmaf(-n, GS_LN2_LO, r);
    float q = GS_EXP_Q_4;
    q = __builtin_fmaf(q, r, GS_EXP_Q_3);
    q = __builtin_fmaf(q, r, GS_EXP_Q_2);
    q = __builtin_fmaf(q, r, GS_EXP_Q_1);
    q = __builtin_fmaf(q, r, GS_EXP_Q_0);
    const float p = __builtin_fmaf(r * r, q, r) + 1.0f;
    const int ni = (int)n;          // in [-150, 128]
    const int n1 = ni >> 1;         // both halves in [-75, 64]
    const int n2 = ni - n1;
    return (p * gs_pow2i(n1)) * gs_pow2i(n2);
}

// gs_expf for x <= 0 without the special-case branches: the argument is
// clamped to -104, where the core path already rounds to +0, so the result is
// bit-identical to gs_expf for every x <= 0 (checked by the test suite).
GS_HD float gs_expf_nonpos(float x)
{
    x = x > -104.0f ? x : -104.0f;
    const float n = __builtin_rintf(x * GS_LOG2E);
    float r = __builtin_fmaf(-n, GS_LN2_HI, x);
    r = __builtin_fmaf(-n, GS_LN2_LO, r);
    float q = GS_EXP_Q_4;
    q = __builtin_fmaf(q, r, GS_EXP_Q_3);
    q = __builtin_fmaf(q, r, GS_EXP_Q_2);
    q = __builtin_fmaf(q, r, GS_EXP_Q_1);
    q = __builtin_fmaf(q, r, GS_EXP_Q_0);
    const float p = __builtin_fmaf(r * r, q, r) + 1.0f;
    const int ni = (int)n;
    const int n1 = ni >> 1;
    const int n2 = ni - n1;
    return (p * gs_pow2i(n1)) * gs_pow2i(n2);
}

// gs_expf on [-86, 0] with fewer operations, for the blend kernels: the
// rounded n = rint(x log2 e) comes from adding 1.5 * 2^23 to the product (the
// float sum rounds to an integer, round-to-nearest-even, exactly as rintf),
// and 2^n is built in one step from that sum's low bits (n >= -124, so p * 2^n
// is a normal float and the two-step scaling of gs_expf gives the same bits).
// Bit-identical to gs_expf for every x in [-86, 0] (checked by the test
// suite); outside that domain the result is unspecified (finite or not, no
// trap).  The blend kernels use exp(power) only where power >= thr =
// gs_alpha_thr(o) >= -86; lanes below thr discard the value.
#define GS_EXP_MAGIC 12582912.0f   // 1.5 * 2^23
GS_HD float gs_expf_blend(float x)
{
    const float t = x * GS_LOG2E + GS_EXP_MAGIC;
    const float n = t - GS_EXP_MAGIC;
    float r = __builtin_fmaf(-n, GS_LN2_HI, x);
    r = __builtin_fmaf(-n, GS_LN2_LO, r);
    float q = GS_EXP_Q_4;
    q = __builtin_fmaf(q, r, GS_EXP_Q_3);
    q = __builtin_fmaf(q, r, GS_EXP_Q_2);
    q = __builtin_fmaf(q, r, GS_EXP_Q_1);
    q = __builtin_fmaf(q, r, GS_EXP_Q_0);
    const float p = __builtin_fmaf(r * r, q, r) + 1.0f;
    // t's bit pattern is 0x4B400000 + n; shifted left by 23 the constant part
    // drops out, leaving the exponent field of 2^n
    return p * gs_as_float((gs_as_uint(t) << 23) + 0x3F800000u);
}

// The blend's skip boundary for opacity o: the smallest float p in [-86, 0]
// with o * gs_expf_blend(p) >= 1/255 in float32 (the test of forward.cu:219-221
// as the blend kernels evaluate it).  gs_expf_blend is non-increasing as its
// argument decreases over all of [-86, 0] (checked exhaustively by the test
// suite) and rounding the product preserves order, so for every power <= 0
//     power >= gs_alpha_thr(o)   <=>   o * gs_expf_blend(power) >= 1/255
// and the forward and backward blends decide with this one comparison.
// The search runs over the bit patterns of the power's magnitude.  Returns +1
// (above every power <= 0) when even power = 0 stays below 1/255 (o < 1/255,
// o <= 0), -86 when every power of the domain passes (o above ~1e35: the
// far tail below -86 is then skipped), and NaN for a NaN opacity (every
// comparison with it is false, so nothing is skipped, as alpha =
// min(0.99, NaN) = 0.99 would not be).
//
// The boundary lies within a few floats of -log(255 o) (the log and the
// exponential are each within about an ulp), so the search starts there and
// steps one float at a time; a boundary further than GS_THR_STEPS floats from
// the start (not expected) falls back to a bisection over all of [-86, 0]
// (31 steps).  Either way the result is the one boundary the comparison
// defines.
#define GS_THR_STEPS 32
GS_HD float gs_logf(float x);
GS_HD float gs_alpha_thr(float o)
{
    const float amin = 1.0f / 255.0f;
    if (o != o) return o;
    if (!(o * gs_expf_blend(0.0f) >= amin)) return 1.0f;
    if (o * gs_expf_blend(-86.0f) >= amin) return -86.0f;
    const uint32_t top = gs_as_uint(86.0f);   // magnitude bits: 0 passes, top does not
    // local search from the estimate (magnitude bits u pass when
    // o * exp(-u) >= 1/255; they pass exactly up to the boundary)
    const float est = gs_logf(255.0f * o);   // -est estimates the boundary power
    uint32_t u0 = est > 0.0f ? gs_as_uint(est) : 0u;
    u0 = u0 < top ? u0 : top - 1u;
    if (o * gs_expf_blend(-gs_as_float(u0)) >= amin) {
        for (int k = 0; k < GS_THR_STEPS; ++k) {   // u0 passes: the last passing above it
            if (!(o * gs_expf_blend(-gs_as_float(u0 + 1u)) >= amin)) return -gs_as_float(u0);
            ++u0;
            if (u0 + 1u >= top) break;
        }
    } else {
        for (int k = 0; k < GS_THR_STEPS && u0 > 0u; ++k) {   // u0 fails: the first passing below it
            --u0;
            if (o * gs_expf_blend(-gs_as_float(u0)) >= amin) return -gs_as_float(u0);
        }
    }
    uint32_t lo = 0u, hi = top;   // bisection: lo passes, hi does not
    while (hi - lo > 1u) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (o * gs_expf_blend(-gs_as_float(mid)) >= amin) lo = mid;
        else hi = mid;
    }
    return -gs_as_float(lo);
}

GS_HD float gs_logf(float x)
{
    if (x != x) return x;
    if (x <= 0.0f) return (x == 0.0f) ? -__builtin_huge_valf() : __builtin_nanf("");
    if (x == __builtin_huge_valf()) return x;
    int eadj = 0;
    if (x < 0x1p-126f) { x *= 0x1p23f; eadj = -23; }
    // Split x = 2^e * m with m in [sqrt(1/2), sqrt(2)).
    const int32_t ix = (int32_t)gs_as_uint(x) - 0x3f3504f3;
    const int e = (ix >> 23) + eadj;
    const float m = gs_as_float((uint32_t)((ix & 0x007fffff) + 0x3f3504f3));
    const float f = m - 1.0f;       // exact (Sterbenz)
    float q = GS_LOG_Q_8;
    q = __builtin_fmaf(q, f, GS_LOG_Q_7);
    q = __builtin_fmaf(q, f, GS_LOG_Q_6);
    q = __builtin_fmaf(q, f, GS_LOG_Q_5);
    q = __builtin_fmaf(q, f, GS_LOG_Q_4);
    q = __builtin_fmaf(q, f, GS_LOG_Q_3);
    q = __builtin_fmaf(q, f, GS_LOG_Q_2);
    q = __builtin_fmaf(q, f, GS_LOG_Q_1);
    q = __builtin_fmaf(q, f, GS_LOG_Q_0);
    const float lp = __builtin_fmaf(f * f, q, f);
    const float ef = (float)e;
    return __builtin_fmaf(ef, GS_LN2_HI, __builtin_fmaf(ef, GS_LN2_LO, lp));
}

GS_HD float gs_atanf_unit(float a)   // a in [0, 1]
{
    const float u = a * a;
    float p = GS_ATAN_P_8;
    p = __builtin_fmaf(p, u, GS_ATAN_P_7);
    p = __builtin_fmaf(p, u, GS_ATAN_P_6);
    p = __builtin_fmaf(p, u, GS_ATAN_P_5);
    p = __builtin_fmaf(p, u, GS_ATAN_P_4);
    p = __builtin_fmaf(p, u, GS_ATAN_P_3);
    p = __builtin_fmaf(p, u, GS_ATAN_P_2);
    p = __builtin_fmaf(p, u, GS_ATAN_P_1);
    p = __builtin_fmaf(p, u, GS_ATAN_P_0);
    return __builtin_fmaf(a * u, p, a);
}

GS_HD float gs_atan2f(float y, float x)
{
    if (x != x || y != y) return x + y;
    const float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y);
    const float mx = ax > ay ? ax : ay;
    const float mn = ax > ay ? ay : ax;
    const float a = (mx == 0.0f) ? 0.0f : mn / mx;
    float t = gs_atanf_unit(a);
    if (ay > ax) t = (GS_PIO2_HI - t) + GS_PIO2_LO;
    if (__builtin_signbit(x)) t = (GS_PI_HI - t) + GS_PI_LO;
    return __builtin_copysignf(t, y);
}

GS_HD float gs_asin_poly(float s, float u)   // asin(s), u = s*s, s in [0, 1/2]
{
    float p = GS_ASIN_S_5;
    p = __builtin_fmaf(p, u, GS_ASIN_S_4);
    p = __builtin_fmaf(p, u, GS_ASIN_S_3);
    p = __builtin_fmaf(p, u, GS_ASIN_S_2);
    p = __builtin_fmaf(p, u, GS_ASIN_S_1);
    p = __builtin_fmaf(p, u, GS_ASIN_S_0);
    return __builtin_fmaf(s * u, p, s);
}

GS_HD float gs_asinf(float x)
{
    const float ax = __builtin_fabsf(x);
    if (!(ax <= 1.0f)) return __builtin_nanf("");
    float r;
    if (ax <= 0.5f) {
        r = gs_asin_poly(ax, ax * ax);
    } else {
        const float z = (1.0f - ax) * 0.5f;   // exact
        const float s = __builtin_sqrtf(z);
        const float as = gs_asin_poly(s, z);
        r = (GS_PIO2_HI - 2.0f * as) + GS_PIO2_LO;
    }
    return __builtin_copysignf(r, x);
}
