// Multi-frequency (CSI) blend: the front-to-back blend of renderCUDA
// (forward.cu:151-249) with GS_CSI_NC complex channels per pixel instead of
// one, and its exact adjoint.  No counterpart in the reference (one channel,
// config.h:15); the emission per subcarrier is defined in gs_csi.h.
//
// alpha and T do not depend on the subcarrier, so the walk, the skips and the
// termination are those of the single-channel blend; only the accumulation
// widens to 104 floats per pixel.  Mapping: one wave per 8x8 quarter tile
// (one pixel per lane, whose 104 values live in VGPRs) or, for the forward,
// one wave per 16x8 half tile (two pixels per lane: each LDS read of an
// entry's emissions then serves two pixels; the forward is bound by the LDS
// return path, not by the FMAs); lists staged 32 entries at a time.  While staging, each lane evaluates the 52 phasors of
// its entry (gs_cycle_frac + gs_sincos2pi, once per entry and wave, not per
// pixel) into LDS: the forward stores the emission e_k = s u_k, the backward
// the phasors u_k.
//
// Backward.  With G_k = dL/dC_k (complex) of a pixel and p_k = conj(G_k) u_k,
// Z = sum_k p_k and Z1 = sum_k k p_k (formed from the top subcarrier down as
// running suffix sums, 6 operations per subcarrier):
//   h = Re(s Z) = sum_k Re(conj(G_k) e_k) plays the colour of the single-
//     channel adjoint (dL/dalpha from h and the running accumulation of h);
//   dL/ds_re = w Re(Z), dL/ds_im = -w Im(Z);
//   dL/dL = w 2 pi Im(s Y), Y = sum_k (f_k / c) p_k = k0 Z + dk Z1.
// Per (entry, wave) the eight g8 values are summed by the transpose reduction
// and dL/dL by a plain wave sum, then added with one atomic each.
#include "gs_blend.h"
#include "gs_csi.h"
#include "gs_cull.h"
#include "gs_device.h"
#include "gs_kernels.h"

namespace gs {

#define CSI_BATCH 32
#define CSI_NC2 (GS_CSI_NC / 2)       // float4 = two subcarriers

template <int PPL>
__global__ void __launch_bounds__(GS_WAVE) k_render_fwd_csi(Dims d, RenderCsiArgs a)
{
    __shared__ float4 s0[CSI_BATCH];
    __shared__ float4 s1[CSI_BATCH];
    __shared__ float4 se[CSI_BATCH][CSI_NC2];
    const Pixels<PPL> pm = map_pixels<PPL>(d, a.order, (int)a.total[1]);
    if (!pm.valid) return;
    if ((uint64_t)a.total[0] > a.Kcap) return;
    bool any_inside = false;
#pragma unroll
    for (int q = 0; q < PPL; ++q) any_inside |= pm.inside[q];
    if (ballot(any_inside) == 0) return;
    const int r = pm.r;
    const int lane = threadIdx.x;
    const uint2 range = a.ranges[(size_t)r * d.T + pm.tile];
    const uint32_t rb = __builtin_amdgcn_readfirstlane(range.x);
    const uint32_t re = __builtin_amdgcn_readfirstlane(range.y);
    const size_t rP = rx_map(d, r).roff;

    // T < 0 marks a finished pixel (as in the single-channel forward)
    float T[PPL], pxf[PPL], pyf[PPL];
    uint32_t last[PPL];
    float acc[PPL][2 * GS_CSI_NC];
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
        T[q] = pm.inside[q] ? 1.0f : -1.0f;
        last[q] = 0;
        pxf[q] = (float)pm.px[q];
        pyf[q] = (float)pm.py[q];
#pragma unroll
        for (int k = 0; k < 2 * GS_CSI_NC; ++k) acc[q][k] = 0.0f;
    }

    for (uint32_t start = rb; start < re; start += CSI_BATCH) {
        bool live = false;
#pragma unroll
        for (int q = 0; q < PPL; ++q) live |= T[q] > 0.0f;
        if (ballot(live) == 0) break;
        const uint32_t e = start + lane;
        bool keep = false;
        if (lane < CSI_BATCH && e < re) {
            const size_t gi = rP + a.plist[e];
            const float4 A = a.splat[2 * gi];
            const float4 B = a.splat[2 * gi + 1];
            s0[lane] = A;
            s1[lane] = B;
            keep = !block_misses(A.x, A.y, A.z, A.w, B.x, B.z, pm.x0, pm.x1, pm.y0, pm.y1);
            if (keep) {
                const float2 S = a.sig[gi];
                const float L = a.plen[gi];
#pragma unroll 2
                for (int k2 = 0; k2 < CSI_NC2; ++k2) {
                    float c0, n0, c1, n1, e0r, e0i, e1r, e1i;
                    gs_sincos2pi(gs_cycle_frac(L, a.csi.khi[2 * k2], a.csi.klo[2 * k2]), &c0, &n0);
                    gs_sincos2pi(gs_cycle_frac(L, a.csi.khi[2 * k2 + 1], a.csi.klo[2 * k2 + 1]), &c1, &n1);
                    gs_csi_emit(S.x, S.y, c0, n0, &e0r, &e0i);
                    gs_csi_emit(S.x, S.y, c1, n1, &e1r, &e1i);
                    se[lane][k2] = make_float4(e0r, e0i, e1r, e1i);
                }
            }
        }
        uint64_t mask = ballot(keep);
        __syncthreads();
        const uint32_t base = start - rb + 1;
        while (mask) {
            const int j = __builtin_ctzll(mask);
            mask &= mask - 1;
            const float4 A = s0[j];
            const float4 B = s1[j];
            float w[PPL];
            bool any = false;
#pragma unroll
            for (int q = 0; q < PPL; ++q) {
                w[q] = 0.0f;
                const float dx = A.x - pxf[q], dy = A.y - pyf[q];
                const float power = pair_power(A.z, A.w, B.x, dx, dy);
                if (power > 0.0f || power < B.z) continue;
                const float alpha = fminf(0.99f, B.y * gs_expf_blend(power));
                if (alpha < 1.0f / 255.0f) continue;
                const float test_T = T[q] * (1.0f - alpha);
                if (test_T < 0.0001f) { T[q] = -fabsf(T[q]); continue; }
                w[q] = alpha * T[q];
                T[q] = test_T;
                last[q] = base + j;
                any = true;
            }
            if (!any) continue;
            // one LDS read of the entry's 52 emissions serves all PPL pixels;
            // a pixel that skipped adds e * 0 = 0 (emissions are finite), which
            // leaves its sums unchanged except for the sign of a zero
#pragma unroll
            for (int k2 = 0; k2 < CSI_NC2; ++k2) {
                const float4 E = se[j][k2];
#pragma unroll
                for (int q = 0; q < PPL; ++q) {
                    acc[q][4 * k2 + 0] = fmaf(E.x, w[q], acc[q][4 * k2 + 0]);
                    acc[q][4 * k2 + 1] = fmaf(E.y, w[q], acc[q][4 * k2 + 1]);
                    acc[q][4 * k2 + 2] = fmaf(E.z, w[q], acc[q][4 * k2 + 2]);
                    acc[q][4 * k2 + 3] = fmaf(E.w, w[q], acc[q][4 * k2 + 3]);
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
        if (!pm.inside[q]) continue;
        const size_t p = (size_t)r * d.W * d.H + (size_t)pm.py[q] * d.W + pm.px[q];
        float4* o = reinterpret_cast<float4*>(a.out + p * (2 * GS_CSI_NC));
#pragma unroll
        for (int k2 = 0; k2 < CSI_NC2; ++k2)
            o[k2] = make_float4(acc[q][4 * k2], acc[q][4 * k2 + 1], acc[q][4 * k2 + 2], acc[q][4 * k2 + 3]);
        a.finalT[p] = fabsf(T[q]);
        a.ncontrib[p] = last[q];
    }
}

__global__ void __launch_bounds__(GS_WAVE) k_render_bwd_csi(Dims d, RenderCsiArgs a)
{
    __shared__ float4 s0[CSI_BATCH];
    __shared__ float4 s1[CSI_BATCH];
    __shared__ float2 ss[CSI_BATCH];
    __shared__ uint32_t sid[CSI_BATCH];
    __shared__ float4 su[CSI_BATCH][CSI_NC2];
    const Pixels<1> pm = map_pixels<1>(d, a.order, (int)a.total[1]);
    if (!pm.valid) return;
    if ((uint64_t)a.total[0] > a.Kcap) return;
    const int r = pm.r;
    const int lane = threadIdx.x;
    const bool inside = pm.inside[0];
    const size_t p = (size_t)r * d.W * d.H + (size_t)pm.py[0] * d.W + pm.px[0];
    float T = inside ? a.finalT[p] : 0.0f;
    const int last = inside ? (int)a.ncontrib[p] : 0;
    float g[2 * GS_CSI_NC];
    if (inside) {
        const float4* gp = reinterpret_cast<const float4*>(a.dout + p * (2 * GS_CSI_NC));
#pragma unroll
        for (int k2 = 0; k2 < CSI_NC2; ++k2) {
            const float4 v = gp[k2];
            g[4 * k2] = v.x; g[4 * k2 + 1] = v.y; g[4 * k2 + 2] = v.z; g[4 * k2 + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 2 * GS_CSI_NC; ++k) g[k] = 0.0f;
    }
    int maxlast = last;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) maxlast = max(maxlast, __shfl_xor(maxlast, off));
    maxlast = __builtin_amdgcn_readfirstlane(maxlast);
    if (maxlast == 0) return;
    const uint2 range = a.ranges[(size_t)r * d.T + pm.tile];
    const uint32_t rb = __builtin_amdgcn_readfirstlane(range.x);
    const size_t rP = rx_map(d, r).roff;
    const float pxf = (float)pm.px[0], pyf = (float)pm.py[0];
    const float two_pi = 6.283185307179586f;

    // accum of backward.cu:579, updated right after each contributor instead
    // of before the next one (the same values, no last-alpha state)
    float acc = 0.0f;
    for (int top = maxlast - 1; top >= 0; top -= CSI_BATCH) {
        // entry j of this batch sits at list position top - j
        const int rel = top - lane;
        bool keep = false;
        if (lane < CSI_BATCH && rel >= 0) {
            const uint32_t id = a.plist[rb + rel];
            const size_t gi = rP + id;
            const float4 A = a.splat[2 * gi];
            const float4 B = a.splat[2 * gi + 1];
            sid[lane] = id;
            s0[lane] = A;
            s1[lane] = B;
            ss[lane] = a.sig[gi];
            keep = !block_misses(A.x, A.y, A.z, A.w, B.x, B.z, pm.x0, pm.x1, pm.y0, pm.y1);
            if (keep) {
                const float L = a.plen[gi];
#pragma unroll 2
                for (int k2 = 0; k2 < CSI_NC2; ++k2) {
                    float c0, n0, c1, n1;
                    gs_sincos2pi(gs_cycle_frac(L, a.csi.khi[2 * k2], a.csi.klo[2 * k2]), &c0, &n0);
                    gs_sincos2pi(gs_cycle_frac(L, a.csi.khi[2 * k2 + 1], a.csi.klo[2 * k2 + 1]), &c1, &n1);
                    su[lane][k2] = make_float4(c0, -n0, c1, -n1);   // u_k = exp(-j 2 pi x)
                }
            }
        }
        uint64_t mask = ballot(keep);
        __syncthreads();
        while (mask) {
            const int j = __builtin_ctzll(mask);
            mask &= mask - 1;
            const float4 A = s0[j];
            const float4 B = s1[j];
            const float2 S = ss[j];
            const float qa = -2.0f * A.z, qb = -A.w, qc = -2.0f * B.x;
            float v[8] = {-0.0f, -0.0f, -0.0f, -0.0f, -0.0f, -0.0f, -0.0f, -0.0f};
            float vL = -0.0f;
            const bool active = (top - j) < last;
            const float dx = A.x - pxf, dy = A.y - pyf;
            const float power = pair_power(A.z, A.w, B.x, dx, dy);
            bool c = active && !(power > 0.0f) && !(power < B.z);
            float G = 0.0f, oG = 0.0f, alpha = 0.0f;
            if (c) {
                G = gs_expf_blend(power);
                oG = B.y * G;
                alpha = fminf(0.99f, oG);
                c = !(alpha < 1.0f / 255.0f);
            }
            if (c) {
                T = T * __builtin_amdgcn_rcpf(1.0f - alpha);
                const float w = alpha * T;
                // subcarriers from the top: S_m = sum_{k >= m} p_k accumulates
                // by FMAs, Z1 = sum_k k p_k = sum_{m >= 1} S_m by adds, Z = S_0
                float zr = 0.0f, zi = 0.0f, z1r = 0.0f, z1i = 0.0f;
#pragma unroll
                for (int k2 = CSI_NC2 - 1; k2 >= 0; --k2) {
                    const float4 U = su[j][k2];
                    const float gr0 = g[4 * k2], gi0 = g[4 * k2 + 1], gr1 = g[4 * k2 + 2], gi1 = g[4 * k2 + 3];
                    zr = fmaf(gr1, U.z, zr); zr = fmaf(gi1, U.w, zr);
                    zi = fmaf(gr1, U.w, zi); zi = fmaf(-gi1, U.z, zi);
                    z1r = z1r + zr; z1i = z1i + zi;
                    zr = fmaf(gr0, U.x, zr); zr = fmaf(gi0, U.y, zr);
                    zi = fmaf(gr0, U.y, zi); zi = fmaf(-gi0, U.x, zi);
                    if (k2 > 0) { z1r = z1r + zr; z1i = z1i + zi; }
                }
                const float h = fmaf(S.x, zr, -(S.y * zi));
                const float eh = h - acc;
                const float dalpha = eh * T;
                acc = fmaf(alpha, eh, acc);
                v[6] = w * zr;
                v[7] = -(w * zi);
                const float yr = fmaf(a.csi.dk, z1r, a.csi.k0 * zr);
                const float yi = fmaf(a.csi.dk, z1i, a.csi.k0 * zi);
                vL = (w * two_pi) * fmaf(S.x, yi, S.y * yr);
                if (!(oG > 0.99f)) {   // clamped alpha does not depend on G or the opacity
                    const float dG = B.y * dalpha;
                    const float gdx = G * dx, gdy = G * dy;
                    v[0] = dG * (-gdx * qa - gdy * qb);
                    v[1] = dG * (-gdy * qc - gdx * qb);
                    v[2] = -0.5f * gdx * dx * dG;
                    v[3] = -gdx * dy * dG;
                    v[4] = -0.5f * gdy * dy * dG;
                    v[5] = G * dalpha;
                }
            }
            if (ballot(c) == 0) continue;
            const float tot = wave_sum8(v, lane);
            const float totL = wave_sum(vL);
            const size_t gi = rP + __builtin_amdgcn_readfirstlane(sid[j]);   // wave-uniform
            if ((lane & 7) == 0) atomicAdd(a.g8 + gi * 8 + (lane >> 3), tot);
            if (lane == 0) atomicAdd(a.gL + gi, totL);
        }
        __syncthreads();
    }
}

hipError_t launch_render_fwd_csi(const Dims& d, const RenderCsiArgs& a, int ppl, hipStream_t s)
{
    if (d.R == 0 || d.T == 0) return hipSuccess;
    if (ppl == 2)
        hipLaunchKernelGGL(k_render_fwd_csi<2>, dim3((unsigned)(order_length(d) * 2)), dim3(GS_WAVE), 0, s, d, a);
    else
        hipLaunchKernelGGL(k_render_fwd_csi<1>, dim3((unsigned)(order_length(d) * 4)), dim3(GS_WAVE), 0, s, d, a);
    return hipGetLastError();
}

hipError_t launch_render_bwd_csi(const Dims& d, const RenderCsiArgs& a, hipStream_t s)
{
    if (d.R == 0 || d.T == 0) return hipSuccess;
    hipLaunchKernelGGL(k_render_bwd_csi, dim3((unsigned)(order_length(d) * 4)), dim3(GS_WAVE), 0, s, d, a);
    return hipGetLastError();
}

}  // namespace gs
