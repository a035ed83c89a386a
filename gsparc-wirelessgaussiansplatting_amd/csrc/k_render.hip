// Front-to-back alpha blend (renderCUDA, forward.cu:151-249) and its adjoint
// (renderCUDA, backward.cu:442-616) for R receivers.
//
// Work unit = one wave64 workgroup per 8x8 quarter of a 16x16 binning tile
// (the 16x16 binning grid is part of the numerics and is kept).  A wave
// stages 64 list entries at a time in LDS and every lane walks them for its
// pixel; waves exit as soon as all 64 of their pixels are saturated, without
// waiting for the other quarters of the tile.  Work-group ids are mapped so
// that the four quarters of a tile run on the same XCD and share its L2.
//
// Per pair the exponential is skipped when power < thr: then
// opacity * exp(power) < 1/255 is guaranteed (thr carries a 1e-3 margin), so
// the decision is the same as evaluating it.
#include "gs_device.h"
#include "gs_kernels.h"

namespace gs {

struct PixelMap {
    int tile, px, py;
    bool valid;
};

__device__ __forceinline__ PixelMap map_pixel(const Dims& d)
{
    PixelMap m;
    const int b = blockIdx.x;
    const int xcd = b & 7, s = b >> 3;
    const int sub = s & 3;
    m.tile = (s >> 2) * 8 + xcd;
    m.valid = m.tile < d.T;
    const int lane = threadIdx.x;
    const int tx = m.valid ? m.tile % d.gx : 0, ty = m.valid ? m.tile / d.gx : 0;
    m.px = tx * GS_BLOCK + (sub & 1) * 8 + (lane & 7);
    m.py = ty * GS_BLOCK + (sub >> 1) * 8 + (lane >> 3);
    return m;
}

__global__ void __launch_bounds__(GS_WAVE) k_render_fwd(Dims d, RenderFwdArgs a)
{
    __shared__ float4 s0[GS_WAVE];
    __shared__ float4 s1[GS_WAVE];
    __shared__ float2 ss[GS_WAVE];
    const PixelMap pm = map_pixel(d);
    if (!pm.valid) return;
    if ((uint64_t)a.total[0] > a.Kcap) return;   // binning overflow, reported by the host
    const bool inside = pm.px < d.W && pm.py < d.H;
    if (ballot(inside) == 0) return;
    const int r = blockIdx.y;
    const int lane = threadIdx.x;
    const uint2 range = a.ranges[(size_t)r * d.T + pm.tile];
    const size_t rP = (size_t)r * d.P;
    const float pxf = (float)pm.px, pyf = (float)pm.py;

    float T = 1.0f, are = 0.0f, aim = 0.0f;
    uint32_t contributor = 0, last = 0;
    bool done = !inside;
    for (uint32_t start = range.x; start < range.y; start += GS_WAVE) {
        if (ballot(!done) == 0) break;
        const uint32_t e = start + lane;
        if (e < range.y) {
            const size_t gi = rP + a.plist[e];
            s0[lane] = a.splat[2 * gi];
            s1[lane] = a.splat[2 * gi + 1];
            ss[lane] = a.sig[gi];
        }
        __syncthreads();
        const int n = (int)min((uint32_t)GS_WAVE, range.y - start);
        for (int j = 0; j < n; ++j) {
            if (done) continue;
            contributor++;
            const float4 A = s0[j];
            const float4 B = s1[j];
            const float dx = A.x - pxf, dy = A.y - pyf;
            const float power = pair_power(A.z, A.w, B.x, dx, dy);
            if (power > 0.0f) continue;
            if (power < B.z) continue;
            const float alpha = fminf(0.99f, B.y * gs_expf(power));
            if (alpha < 1.0f / 255.0f) continue;
            const float test_T = T * (1.0f - alpha);
            if (test_T < 0.0001f) { done = true; continue; }
            const float2 S = ss[j];
            const float w = alpha * T;
            are = fmaf(S.x, w, are);
            aim = fmaf(S.y, w, aim);
            T = test_T;
            last = contributor;
        }
        __syncthreads();
    }
    if (a.pairs) {
        unsigned long long c = contributor;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off);
        if (lane == 0) atomicAdd(a.pairs, c);
    }
    if (inside) {
        const size_t p = (size_t)r * d.W * d.H + (size_t)pm.py * d.W + pm.px;
        a.out_re[p] = are;
        a.out_im[p] = aim;
        a.finalT[p] = T;
        a.ncontrib[p] = last;
    }
}

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

template <int MODE>
__global__ void __launch_bounds__(GS_WAVE) k_render_bwd(Dims d, RenderBwdArgs a)
{
    __shared__ float4 s0[GS_WAVE];
    __shared__ float4 s1[GS_WAVE];
    __shared__ float2 ss[GS_WAVE];
    __shared__ uint32_t sid[GS_WAVE];
    const PixelMap pm = map_pixel(d);
    if (!pm.valid) return;
    if ((uint64_t)a.total[0] > a.Kcap) return;
    const bool inside = pm.px < d.W && pm.py < d.H;
    const int r = blockIdx.y;
    const int lane = threadIdx.x;
    const size_t p = (size_t)r * d.W * d.H + (size_t)pm.py * d.W + pm.px;
    float T = inside ? a.finalT[p] : 0.0f;
    const int last = inside ? (int)a.ncontrib[p] : 0;
    int maxlast = last;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) maxlast = max(maxlast, __shfl_xor(maxlast, off));
    if (maxlast == 0) return;
    const float gre = inside ? a.dout_re[p] : 0.0f;
    const float gim = inside ? a.dout_im[p] : 0.0f;
    const uint2 range = a.ranges[(size_t)r * d.T + pm.tile];
    const size_t rP = (size_t)r * d.P;
    const float pxf = (float)pm.px, pyf = (float)pm.py;

    float acc_re = 0.0f, acc_im = 0.0f, last_alpha = 0.0f, last_re = 0.0f, last_im = 0.0f;
    for (int top = maxlast - 1; top >= 0; top -= GS_WAVE) {
        const int rel = top - lane;
        if (rel >= 0) {
            const uint32_t id = a.plist[range.x + rel];
            const size_t gi = rP + id;
            sid[lane] = id;
            s0[lane] = a.splat[2 * gi];
            s1[lane] = a.splat[2 * gi + 1];
            ss[lane] = a.sig[gi];
        }
        __syncthreads();
        const int n = min(GS_WAVE, top + 1);
        for (int j = 0; j < n; ++j) {
            const bool active = (top - j) < last;
            const float4 A = s0[j];
            const float4 B = s1[j];
            const float dx = A.x - pxf, dy = A.y - pyf;
            const float power = pair_power(A.z, A.w, B.x, dx, dy);
            bool c = active && !(power > 0.0f) && !(power < B.z);
            float G = 0.0f, oG = 0.0f, alpha = 0.0f;
            if (c) {
                G = gs_expf(power);
                oG = B.y * G;
                alpha = fminf(0.99f, oG);
                c = !(alpha < 1.0f / 255.0f);
            }
            float v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (c) {
                T = T / (1.0f - alpha);
                const float w = alpha * T;
                const float2 S = ss[j];
                const float sr = S.x;
                const float si = (MODE == GRAD_REFERENCE) ? S.x : S.y;
                acc_re = last_alpha * last_re + (1.0f - last_alpha) * acc_re;
                acc_im = last_alpha * last_im + (1.0f - last_alpha) * acc_im;
                last_re = sr;
                last_im = si;
                float dalpha = (sr - acc_re) * gre;
                if (MODE == GRAD_EXACT) dalpha += (si - acc_im) * gim;
                dalpha *= T;
                last_alpha = alpha;
                v[6] = w * gre;
                v[7] = w * gim;
                if (!(MODE == GRAD_EXACT && oG > 0.99f)) {
                    const float qa = -2.0f * A.z, qb = -A.w, qc = -2.0f * B.x;
                    const float dG = B.y * dalpha;
                    const float gdx = G * dx, gdy = G * dy;
                    const float dG_dx = -gdx * qa - gdy * qb;
                    const float dG_dy = -gdy * qc - gdx * qb;
                    v[0] = dG * dG_dx;
                    v[1] = dG * dG_dy;
                    v[2] = -0.5f * gdx * dx * dG;
                    v[3] = -gdx * dy * dG;
                    v[4] = -0.5f * gdy * dy * dG;
                    v[5] = G * dalpha;
                }
            }
            if (ballot(c) == 0) continue;
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = wave_sum(v[k]);
            if (lane < 8) {
                float x = v[0];
#pragma unroll
                for (int k = 1; k < 8; ++k) x = (lane == k) ? v[k] : x;
                atomicAdd(a.g8 + (rP + sid[j]) * 8 + lane, x);
            }
        }
        __syncthreads();
    }
}

static inline dim3 render_grid(const Dims& d)
{
    const int tiles8 = (d.T + 7) / 8 * 8;
    return dim3(tiles8 * 4, d.R);
}

hipError_t launch_render_fwd(const Dims& d, const RenderFwdArgs& a, hipStream_t s)
{
    if (d.R == 0 || d.T == 0) return hipSuccess;
    hipLaunchKernelGGL(k_render_fwd, render_grid(d), dim3(GS_WAVE), 0, s, d, a);
    return hipGetLastError();
}

hipError_t launch_render_bwd(const Dims& d, const RenderBwdArgs& a, hipStream_t s)
{
    if (d.R == 0 || d.T == 0) return hipSuccess;
    if (a.mode == GRAD_REFERENCE)
        hipLaunchKernelGGL(k_render_bwd<GRAD_REFERENCE>, render_grid(d), dim3(GS_WAVE), 0, s, d, a);
    else
        hipLaunchKernelGGL(k_render_bwd<GRAD_EXACT>, render_grid(d), dim3(GS_WAVE), 0, s, d, a);
    return hipGetLastError();
}

}  // namespace gs
