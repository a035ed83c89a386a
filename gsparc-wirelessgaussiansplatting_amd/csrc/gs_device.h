// gs_device.h -- constants, record layouts and per-Gaussian geometry shared by
// the gfx950 rasterizer kernels.
//
// Layout in HBM (r = receiver, g = Gaussian, P Gaussians, R receivers,
// T = tiles per map, N = H*W pixels):
//   splat[r*P + g]  = 2 x float4 {px, py, ha, nb} {hc, opacity, thr, 0}
//                     ha = -qa/2, nb = -qb, hc = -qc/2 (conic (qa, qb, qc)),
//                     thr = gs_alpha_thr(o): a pair has alpha >= 1/255
//                     exactly when power >= thr, so both blends decide by
//                     this comparison and skip the exponential below it.
//   sig[r*P + g]    = float2 {real, imag} emission
//   depth[r*P + g]  = float distance to the receiver (sort key)
//   radii[r*P + g]  = int, 0 = culled
//   segcnt[r*T + t] = instances of tile t of receiver r; segbase = its
//                     exclusive prefix over (r, t); cursor = scatter fill count
//   keys/vals[K]    = depth bits / Gaussian index, grouped by (receiver, tile)
//   plist[K]        = Gaussian index, sorted by (receiver, tile, depth, idx)
//   finalT[r*N + p], ncontrib[r*N + p] = per-pixel blend state
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gs_math.h"

#define GS_BLOCK 16                       // binning tile edge (config.h:16-17)
#define GS_WAVE 64
#define GS_PRE_THREADS 256                // Gaussians per preprocess block
#define GS_ONE_OVER_PI_F 0.318309886183790671538f   // M_1_PIf32
#define GS_TWO_OVER_PI_F 0.636619772367581343076f   // M_2_PIf32

namespace gs {

enum GradMode { GRAD_EXACT = 0, GRAD_REFERENCE = 1 };

struct GeoOut {
    float t[3];
    float r;
    float px, py;
    float J[2][3];
};

__device__ __forceinline__ void view_xform(float x, float y, float z, const float* m, float* t)
{
    t[0] = x * m[0] + y * m[4] + z * m[8] + m[12];
    t[1] = x * m[1] + y * m[5] + z * m[9] + m[13];
    t[2] = x * m[2] + y * m[6] + z * m[10] + m[14];
}

// cov3D = Rq diag((mod s)^2) Rq^T, entries xx xy xz yy yz zz
// (forward.cu:124-149).
__device__ __forceinline__ void cov3d(const float* s, float mod, const float* q, float* c)
{
    const float sx = mod * s[0], sy = mod * s[1], sz = mod * s[2];
    const float r = q[0], x = q[1], y = q[2], z = q[3];
    const float Rq[3][3] = {
        {1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y - r * z), 2.0f * (x * z + r * y)},
        {2.0f * (x * y + r * z), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z - r * x)},
        {2.0f * (x * z - r * y), 2.0f * (y * z + r * x), 1.0f - 2.0f * (x * x + y * y)}};
    const float s2[3] = {sx * sx, sy * sy, sz * sz};
    const int ij[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
#pragma unroll
    for (int e = 0; e < 6; ++e) {
        const int i = ij[e][0], j = ij[e][1];
        c[e] = Rq[i][0] * s2[0] * Rq[j][0] + Rq[i][1] * s2[1] * Rq[j][1] + Rq[i][2] * s2[2] * Rq[j][2];
    }
}

// Equirect projection and its Jacobian (forward.cu:82-97, auxiliary.h:10-13,
// 101-113); px, py via ndc2Pix in double as the reference does.
__device__ __forceinline__ void project(const float* t, float r, int W, int H, GeoOut& g)
{
    const float inv_r = 1.0f / (r + 0.0000001f);
    const float lon = gs_atan2f(t[0], t[2]);
    const float lat = gs_asinf(t[1] * inv_r);
    const float sx = lon * GS_ONE_OVER_PI_F;
    const float sy = lat * GS_TWO_OVER_PI_F;
    g.px = (float)((((double)sx + 1.0) * W - 1.0) * 0.5);
    g.py = (float)((((double)sy + 1.0) * H - 1.0) * 0.5);
    const float rxz2 = t[0] * t[0] + t[2] * t[2];
    const float rxz2_inv = 1.0f / (rxz2 + 1e-7f);
    const float rxz = sqrtf(rxz2);
    const float rxz_inv = 1.0f / (rxz + 1e-7f);
    const float r2 = rxz2 + t[1] * t[1];
    const float r2_inv = 1.0f / (r2 + 1e-7f);
    const float Wd = (float)W * 0.5f * GS_ONE_OVER_PI_F;
    const float Hd = (float)H * GS_ONE_OVER_PI_F;
    g.J[0][0] = Wd * t[2] * rxz2_inv;
    g.J[0][1] = 0.0f;
    g.J[0][2] = -Wd * t[0] * rxz2_inv;
    g.J[1][0] = -Hd * t[0] * t[1] * rxz_inv * r2_inv;
    g.J[1][1] = Hd * rxz * r2_inv;
    g.J[1][2] = -Hd * t[2] * t[1] * rxz_inv * r2_inv;
}

// (J Rv) Sigma (J Rv)^T + 0.3 I (forward.cu:99-121).  M2 receives J Rv.
__device__ __forceinline__ void cov2d(const float J[2][3], const float* view, const float* c, float M2[2][3],
                                      float& A, float& B, float& C)
{
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            M2[a][k] = J[a][0] * view[k * 4 + 0] + J[a][1] * view[k * 4 + 1] + J[a][2] * view[k * 4 + 2];
    const float S[3][3] = {{c[0], c[1], c[2]}, {c[1], c[3], c[4]}, {c[2], c[4], c[5]}};
    float v0[3], v1[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        v0[i] = S[i][0] * M2[0][0] + S[i][1] * M2[0][1] + S[i][2] * M2[0][2];
        v1[i] = S[i][0] * M2[1][0] + S[i][1] * M2[1][1] + S[i][2] * M2[1][2];
    }
    A = M2[0][0] * v0[0] + M2[0][1] * v0[1] + M2[0][2] * v0[2] + 0.3f;
    B = M2[1][0] * v0[0] + M2[1][1] * v0[1] + M2[1][2] * v0[2];
    C = M2[1][0] * v1[0] + M2[1][1] * v1[1] + M2[1][2] * v1[2] + 0.3f;
}

// getRect (auxiliary.h:14-24): tiles [x0,x1) x [y0,y1).
__device__ __forceinline__ void get_rect(float px, float py, int radius, int gx, int gy, int& x0, int& y0, int& x1,
                                         int& y1)
{
    const float rr = (float)radius;
    x0 = min(gx, max(0, (int)((px - rr) / (float)GS_BLOCK)));
    y0 = min(gy, max(0, (int)((py - rr) / (float)GS_BLOCK)));
    x1 = min(gx, max(0, (int)(((px + rr) + (float)(GS_BLOCK - 1)) / (float)GS_BLOCK)));
    y1 = min(gy, max(0, (int)(((py + rr) + (float)(GS_BLOCK - 1)) / (float)GS_BLOCK)));
}

// power = -0.5 (qa dx^2 + qc dy^2) - qb dx dy (forward.cu:215) in the
// evaluation order shared with the CPU restatement.
__device__ __forceinline__ float pair_power(float ha, float nb, float hc, float dx, float dy)
{
    const float e = fmaf(nb, dy, ha * dx);
    return fmaf(dx, e, (hc * dy) * dy);
}

// ---- wave helpers (wave64) ----
// The mask of active lanes where p holds.  The builtin takes p as a lane
// mask, so a predicate that is already a compare or a combination of
// compares costs no vector instruction (HIP's __ballot goes through an
// integer compare of p).
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

// Inclusive prefix over the wave's 64 lanes by DPP, no LDS round trip:
// row_shr 1, 2, 4, 8 scan each row of 16 lanes (lanes without a source keep
// the identity), then row_bcast 15 carries rows 0 and 2 into rows 1 and 3 and
// row_bcast 31 carries lane 31 into rows 2 and 3.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_or_identity(int identity, int v)
{
    return __builtin_amdgcn_update_dpp(identity, v, CTRL, ROWS, 0xf, false);
}

// exp(x) as the hardware 2^t (v_exp_f32) of t = x * log2(e), for the blend
// backward, where G = exp(power) only feeds gradients: the forward's
// decisions come in as the contributor count and final T, and the skip test
// is the forward's own power >= thr comparison (gs_alpha_thr), never a test
// of this value.  For power in
// [thr, 0] with opacity <= 1 (|t| < 8.1) the relative error is below 4.5e-7
// (the rounding of t, at most 8.1 * 2^-24 * ln 2, plus 1 ulp of v_exp_f32),
// far inside the 1e-4 gradient tolerance.  Measured against the correctly
// rounded 2^t on all of [-8.125, 0], v_exp_f32 is off by one ulp for 2.1% of
// the inputs and never by more.
__device__ __forceinline__ float gs_expf_grad(float x)
{
    return __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
}

__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t x)
{
    int v = (int)x;
    v += dpp_or_identity<0x111, 0xf>(0, v);
    v += dpp_or_identity<0x112, 0xf>(0, v);
    v += dpp_or_identity<0x114, 0xf>(0, v);
    v += dpp_or_identity<0x118, 0xf>(0, v);
    v += dpp_or_identity<0x142, 0xa>(0, v);
    v += dpp_or_identity<0x143, 0xc>(0, v);
    return (uint32_t)v;
}

// Inclusive prefix maximum of values >= -1 (identity -1).
__device__ __forceinline__ int wave_incl_max(int v)
{
    v = max(v, dpp_or_identity<0x111, 0xf>(-1, v));
    v = max(v, dpp_or_identity<0x112, 0xf>(-1, v));
    v = max(v, dpp_or_identity<0x114, 0xf>(-1, v));
    v = max(v, dpp_or_identity<0x118, 0xf>(-1, v));
    v = max(v, dpp_or_identity<0x142, 0xa>(-1, v));
    v = max(v, dpp_or_identity<0x143, 0xc>(-1, v));
    return v;
}

}  // namespace gs
