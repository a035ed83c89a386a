// gs_csi.h -- multi-frequency (CSI) emission, shared by the gfx950 kernels and
// the host code that checks them.  No counterpart in the reference, which
// renders one complex channel (config.h:15, NUM_CHANNELS 1).
//
// Definition.  Subcarrier k = 0..GS_CSI_NC-1 has frequency
//     f_k = f0 + k * df            (default f0 = 3.5 GHz - 25.5 * 312.5 kHz, df = 312.5 kHz)
// A (receiver, Gaussian) record with complex amplitude s (the per-Gaussian
// MLP output, signal_real + j signal_imag) and path length
//     L = |tx - mu| + |mu - rx|    (transmitter -> Gaussian -> receiver)
// emits on subcarrier k
//     e_k = s * exp(-j 2 pi f_k L / c),
// and a pixel's CSI is C_k = sum_i w_i e_ik with the usual blend weights
// w_i = alpha_i T_i (alpha and T do not depend on the subcarrier).
//
// Evaluation in float32.  The phase is computed as a cycle count n_k = L *
// (f_k / c).  f_k / c is split on the host into khi + klo (float32 each); with
// p = L * khi rounded, e = fma(L, khi, -p) is its exact rounding error, so
//     x = (p - rint(p)) + fma(L, klo, e)
// is frac(n_k) to a few ulp of |x| <= 1/2, whatever the size of n_k (~200
// cycles at 20 m); a plain float32 product would lose ~1e-5 cycles there.
// (cos 2 pi x, sin 2 pi x) come from gs_sincos2pi: quadrant reduction
// y = x - q/4 (exact), |y| <= 1/8, then near-minimax polynomials in y^2
// (coefficients fitted in double and rounded to float32; max abs error
// ~9e-8, checked by the test suite against double precision).
#pragma once

#include "gs_math.h"

#define GS_CSI_NC 52                                   // subcarriers
#define GS_CSI_F0 (3.5e9 - 25.5 * 312.5e3)             // Hz
#define GS_CSI_DF 312.5e3                              // Hz
#define GS_LIGHT_SPEED 299792458.0                     // m/s

// f_k / c split for the float32 cycle count; k0 and dk are f0 / c and df / c
// rounded to float32 (for the path-length gradient, which is linear in f_k);
// dklo = df / c - dk, for the cycle count of the subcarrier spacing.
struct GsCsi {
    float khi[GS_CSI_NC];
    float klo[GS_CSI_NC];
    float k0, dk;
    float dklo;
};

GS_HD GsCsi gs_csi_make(double f0_hz, double df_hz, double c_mps)
{
    GsCsi t;
    for (int k = 0; k < GS_CSI_NC; ++k) {
        const double K = (f0_hz + (double)k * df_hz) / c_mps;
        t.khi[k] = (float)K;
        t.klo[k] = (float)(K - (double)t.khi[k]);
    }
    t.k0 = (float)(f0_hz / c_mps);
    t.dk = (float)(df_hz / c_mps);
    t.dklo = (float)(df_hz / c_mps - (double)t.dk);
    return t;
}

// frac(L * K) with K = khi + klo, in [-1/2, 1/2] up to rounding
GS_HD float gs_cycle_frac(float L, float khi, float klo)
{
    const float p = L * khi;
    const float e = __builtin_fmaf(L, khi, -p);
    const float t = __builtin_fmaf(L, klo, e);
    const float n = __builtin_rintf(p);
    return (p - n) + t;
}

#define GS_SIN2PI_0 6.283185482e+00f
#define GS_SIN2PI_1 -4.134170151e+01f
#define GS_SIN2PI_2 8.160517883e+01f
#define GS_SIN2PI_3 -7.669579315e+01f
#define GS_SIN2PI_4 4.141379929e+01f
#define GS_COS2PI_1 -1.973920822e+01f
#define GS_COS2PI_2 6.493931580e+01f
#define GS_COS2PI_3 -8.544281769e+01f
#define GS_COS2PI_4 5.921903992e+01f

// (cos 2 pi x, sin 2 pi x) for |x| <= 1/2 (a little beyond is fine)
GS_HD void gs_sincos2pi(float x, float* c, float* s)
{
    const float q = __builtin_rintf(4.0f * x);
    const float y = __builtin_fmaf(q, -0.25f, x);
    const float z = y * y;
    float ps = __builtin_fmaf(GS_SIN2PI_4, z, GS_SIN2PI_3);
    ps = __builtin_fmaf(ps, z, GS_SIN2PI_2);
    ps = __builtin_fmaf(ps, z, GS_SIN2PI_1);
    ps = __builtin_fmaf(ps, z, GS_SIN2PI_0);
    const float sy = y * ps;
    float pc = __builtin_fmaf(GS_COS2PI_4, z, GS_COS2PI_3);
    pc = __builtin_fmaf(pc, z, GS_COS2PI_2);
    pc = __builtin_fmaf(pc, z, GS_COS2PI_1);
    const float cy = __builtin_fmaf(pc, z, 1.0f);
    const int iq = ((int)q) & 3;
    // rotate by q quarter turns
    if (iq == 0) { *c = cy; *s = sy; }
    else if (iq == 1) { *c = -sy; *s = cy; }
    else if (iq == 2) { *c = -cy; *s = -sy; }
    else { *c = sy; *s = -cy; }
}

// Emission of subcarrier k: s * exp(-j 2 pi x), with (cs, sn) = (cos, sin) 2 pi x.
GS_HD void gs_csi_emit(float sre, float sim, float cs, float sn, float* er, float* ei)
{
    *er = __builtin_fmaf(sre, cs, sim * sn);
    *ei = __builtin_fmaf(sim, cs, -(sre * sn));
}

// Transmitter -> Gaussian -> receiver path length.
GS_HD float gs_path_length(float mx, float my, float mz, const float* tx, const float* rx)
{
    const float ax = mx - tx[0], ay = my - tx[1], az = mz - tx[2];
    const float bx = mx - rx[0], by = my - rx[1], bz = mz - rx[2];
    const float da = __builtin_sqrtf(ax * ax + ay * ay + az * az);
    const float db = __builtin_sqrtf(bx * bx + by * by + bz * bz);
    return da + db;
}
