// gs_cull.h -- wave-level culling test of a tile-list entry against a block
// of pixel centres, shared by the gfx950 blend kernels and the host test that
// checks it against brute force.
//
// block_misses() returns true only when every pixel centre (px, py) of the
// block [x0,x1] x [y0,y1] is certain to evaluate
//     power = __builtin_fmaf(dx, __builtin_fmaf(nb, dy, ha*dx), (hc*dy)*dy) < thr
// in float32 (dx = mx - px, dy = my - py), so that skipping the entry for the
// whole block changes no result bit of the blend.
//
// power(d) = ha dx^2 + nb dx dy + hc dy^2 is concave for a valid conic
// (ha, hc < 0), so its maximum over the rectangle of offsets is 0 when the
// rectangle holds the origin and otherwise lies on an edge, where it is a 1-D
// concave quadratic maximised in closed form and clamped.  Rounding: the
// per-pixel float evaluation (with the rounded offsets) exceeds the exact
// value by at most ~5u * S, and the float maximum here undershoots the exact
// one by at most ~4u * S (an inexact arg-max only lowers it by |hc| d^2, far
// below that; a non-finite arg-max factor is never culled), where u = 2^-24 and S bounds |ha| dx^2 + |nb||dx dy| + |hc| dy^2
// over the block.  The test keeps a margin of 4e-6 * S (~67u * S).  Anything
// non-finite or non-concave is never culled.
#pragma once

#include "gs_math.h"

GS_HD float gs_rcp_approx(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

GS_HD float gs_clampf(float v, float lo, float hi) { return __builtin_fminf(__builtin_fmaxf(v, lo), hi); }

// ha u^2 + nb u w + hc w^2 for the point (u, w) of an edge
GS_HD float gs_quad(float ha, float nb, float hc, float u, float w)
{
    return __builtin_fmaf(ha * u, u, __builtin_fmaf(nb * u, w, hc * w * w));
}

GS_HD bool block_misses(float mx, float my, float ha, float nb, float hc, float thr, float x0, float x1, float y0,
                        float y1, float margin = 4.0e-6f)
{
    if (!(ha < 0.0f && hc < 0.0f)) return false;
    const float dxl = mx - x1, dxh = mx - x0, dyl = my - y1, dyh = my - y0;
    const float DX = __builtin_fmaxf(__builtin_fabsf(dxl), __builtin_fabsf(dxh));
    const float DY = __builtin_fmaxf(__builtin_fabsf(dyl), __builtin_fabsf(dyh));
    const float smax = gs_quad(-ha, __builtin_fabsf(nb), -hc, DX, DY);
    float pmax = 0.0f;
    if (!(dxl <= 0.0f && dxh >= 0.0f && dyl <= 0.0f && dyh >= 0.0f)) {
        // on an edge x = X the arg-max is y = ky X; on an edge y = Y it is x = kx Y
        const float ky = -0.5f * nb * gs_rcp_approx(hc);
        const float kx = -0.5f * nb * gs_rcp_approx(ha);
        if (!(__builtin_fabsf(ky) <= 3.0e38f && __builtin_fabsf(kx) <= 3.0e38f)) return false;
        const float e0 = gs_quad(ha, nb, hc, dxl, gs_clampf(ky * dxl, dyl, dyh));
        const float e1 = gs_quad(ha, nb, hc, dxh, gs_clampf(ky * dxh, dyl, dyh));
        const float e2 = gs_quad(ha, nb, hc, gs_clampf(kx * dyl, dxl, dxh), dyl);
        const float e3 = gs_quad(ha, nb, hc, gs_clampf(kx * dyh, dxl, dxh), dyh);
        if (!(e0 == e0 && e1 == e1 && e2 == e2 && e3 == e3)) return false;
        pmax = __builtin_fmaxf(__builtin_fmaxf(e0, e1), __builtin_fmaxf(e2, e3));
    }
    return __builtin_fmaf(margin, smax, pmax) < thr;
}
