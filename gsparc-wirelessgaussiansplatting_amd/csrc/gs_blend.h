// gs_blend.h -- pieces shared by the blend kernels (k_render.hip,
// k_render_csi.hip): the mapping of workgroups to (receiver, tile, quarter)
// pixel blocks, and the wave-level reductions of per-entry gradients.
#pragma once

#include "gs_device.h"
#include "gs_kernels.h"

namespace gs {

template <int PPL>
struct Pixels {
    int r, tile;
    bool valid;
    int px[PPL], py[PPL];
    bool inside[PPL];
    float x0, x1, y0, y1;   // pixel-centre bounds of the wave's block (inclusive)
    float qx[PPL], qy[PPL]; // first pixel centre of each lane pixel's 8x8 quarter
};

// Workgroup b runs on XCD b & 7 (workgroups are dealt to the XCDs round
// robin).  Its wave takes slot s = b >> 3 of that XCD: the work item is
// entry s / WPT of the XCD's order list (ORDER_XCD_*) or of the global list
// dealt round robin (ORDER_GLOBAL); the WPT waves of an item share the XCD
// and its L2.  Items come longest list first, so that they do not start last
// and set the kernel's end.
template <int PPL>
__device__ __forceinline__ Pixels<PPL> map_pixels(const Dims& d, const uint32_t* order, int order_mode)
{
    constexpr int WPT = 4 / PPL;   // waves per tile
    Pixels<PPL> m;
    const int b = blockIdx.x;
    const int xcd = b & 7, s = b >> 3;
    const int sub = s % WPT;
    const size_t n8 = (size_t)(d.R + 7) / 8 * d.T;
    const size_t idx = order_mode == ORDER_GLOBAL ? (size_t)(s / WPT) * 8 + xcd : xcd * n8 + s / WPT;
    const uint32_t rt = idx < 8 * n8 ? order[idx] : 0xffffffffu;
    m.valid = rt != 0xffffffffu;
    m.r = (int)(rt / (uint32_t)d.T);
    m.tile = (int)(rt % (uint32_t)d.T);
    const int lane = threadIdx.x;
    const int tx = m.valid ? m.tile % d.gx : 0, ty = m.valid ? m.tile / d.gx : 0;
    int qx0 = 1, qx1 = 0, qy0 = 1, qy1 = 0;
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int q = sub * PPL + k;   // quarter of the tile
        m.px[k] = tx * GS_BLOCK + (q & 1) * 8 + (lane & 7);
        m.py[k] = ty * GS_BLOCK + (q >> 1) * 8 + (lane >> 3);
        m.inside[k] = m.px[k] < d.W && m.py[k] < d.H;
        m.qx[k] = (float)(tx * GS_BLOCK + (q & 1) * 8);
        m.qy[k] = (float)(ty * GS_BLOCK + (q >> 1) * 8);
        qx0 = min(qx0, q & 1); qx1 = max(qx1, q & 1);
        qy0 = min(qy0, q >> 1); qy1 = max(qy1, q >> 1);
    }
    m.x0 = (float)(tx * GS_BLOCK + qx0 * 8);
    m.x1 = (float)(tx * GS_BLOCK + qx1 * 8 + 7);
    m.y0 = (float)(ty * GS_BLOCK + qy0 * 8);
    m.y1 = (float)(ty * GS_BLOCK + qy1 * 8 + 7);
    return m;
}

// The tile lists overflowed their capacity (K > Kcap, so the placement
// wrote nothing): the forward blends do not run.  They leave defined
// outputs instead, written by the kernel itself (nothing to clear in a
// captured graph): NaN spectra, so that a loss computed from them cannot
// pass for a real one, transmittance 1 and no contributors, so that the
// backward (which returns without reading them) has a consistent state.
// out: nch floats per pixel; out2 (may be null) one more.
template <int PPL>
__device__ __forceinline__ void overflow_outputs(const Dims& d, const Pixels<PPL>& pm, float* out, float* out2,
                                                 int nch, float* finalT, uint32_t* ncontrib)
{
    const float qnan = __builtin_nanf("");
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        if (!pm.inside[k]) continue;
        const size_t p = (size_t)pm.r * d.W * d.H + (size_t)pm.py[k] * d.W + pm.px[k];
        for (int c = 0; c < nch; ++c) out[p * nch + c] = qnan;
        if (out2) out2[p] = qnan;
        finalT[p] = 1.0f;
        ncontrib[p] = 0u;
    }
}

template <int CTRL>
__device__ __forceinline__ float dpp_move(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

// Lane-pair sums through the gfx950 lane-swap instructions (VALU, no LDS
// round trip).  v_permlane32_swap exchanges lanes 32..63 of its first operand
// with lanes 0..31 of its second, so after swap(a, b) the two results add up
// to a[l] + a[l + 32] in lanes 0..31 and b[l - 32] + b[l] in lanes 32..63;
// v_permlane16_swap does the same across 16 lanes within each half.
__device__ __forceinline__ float swap_sum32(float a, float b)
{
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

__device__ __forceinline__ float swap_sum16(float a, float b)
{
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// Sums each of v[0..7] over the 64 lanes; lane 8*i returns the total of v[i]
// (every lane of the 8-lane group i holds it).  Transpose reduction: the
// exchanges across 32 and 16 lanes (lane swaps) and 8 lanes (DPP row_ror:8)
// each halve the number of values a lane carries (4, 2, 1), then an 8-lane
// sum: quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror.
__device__ __forceinline__ float wave_sum8(const float (&v)[8], int lane)
{
    const bool b3 = lane & 8;
    float w[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) w[i] = swap_sum32(v[i], v[i + 4]);   // lanes 0..31: v[i]; 32..63: v[i+4]
    float u[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) u[i] = swap_sum16(w[i], w[i + 2]);   // lanes with bit 4 clear: w[i]
    float x = (b3 ? u[1] : u[0]) + dpp_move<0x128>(b3 ? u[0] : u[1]);   // row_ror:8 = lane ^ 8
    x = x + dpp_move<0xb1>(x);
    x = x + dpp_move<0x4e>(x);
    x = x + dpp_move<0x141>(x);
    return x;
}

// Sum of x over the 64 lanes (every lane gets it).
__device__ __forceinline__ float wave_sum(float x)
{
    x = swap_sum32(x, x);
    x = swap_sum16(x, x);
    x = x + dpp_move<0x128>(x);
    x = x + dpp_move<0xb1>(x);
    x = x + dpp_move<0x4e>(x);
    x = x + dpp_move<0x141>(x);
    return x;
}

}  // namespace gs
