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

// (cos 2 pi x, sin 2 pi x) of the forward's phase factors: the hardware
// v_cos_f32 / v_sin_f32 (argument in turns) of the exact cycle fraction.
// The phases scale the emitted signal only; no decision of the blend depends
// on them, and the maps stay within 1e-5 of the oracle's peak.
__device__ __forceinline__ void csi_phase(float x, float* c, float* s)
{
    *c = __builtin_amdgcn_cosf(x);
    *s = __builtin_amdgcn_sinf(x);
}

// csi_phase over an array, for its accuracy check
__global__ void __launch_bounds__(256) k_phase_factors(size_t n, const float* x, float* c, float* s)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        float cv, sv;
        csi_phase(x[i], &cv, &sv);
        c[i] = cv;
        s[i] = sv;
    }
}

hipError_t launch_phase_factors(size_t n, const float* x, float* c, float* s, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_phase_factors, dim3((unsigned)std::min<size_t>((n + 255) / 256, 65536)), dim3(256), 0, st,
                       n, x, c, s);
    return hipGetLastError();
}

// Phasors u_k = exp(-j 2 pi L f_k / c) of one half of the subcarriers,
// k = h*NC/2 .. h*NC/2 + NC/2 - 1, as NC/4 float4 (two subcarriers each,
// handed to put(m, value)): the exact phase at the half's first subcarrier,
// then the complex recurrence u_{k+1} = u_k rho with rho = exp(-j 2 pi L df / c).
// Both start from exact cycle fractions; the first phasor takes the hardware
// sin/cos, rho the polynomial gs_sincos2pi, because rho's error accumulates
// over the steps and the polynomial is the more accurate of the two at rho's
// small angles.  Four operations per subcarrier instead of a cycle fraction
// and two transcendentals; the error grows linearly over the 25 steps (the
// measured bounds are in the phasor accuracy test).  Used where the phasors
// weight gradients only.
static_assert(GS_CSI_NC % 4 == 0, "two halves of whole float4 pairs");
template <typename Put>
__device__ __forceinline__ void csi_phasor_half(float L, int h, const GsCsi& t, Put&& put)
{
    const float x0 = gs_cycle_frac(L, h ? t.khi[GS_CSI_NC / 2] : t.khi[0], h ? t.klo[GS_CSI_NC / 2] : t.klo[0]);
    const float xs = gs_cycle_frac(L, t.dk, t.dklo);
    float ur = __builtin_amdgcn_cosf(x0), ui = -__builtin_amdgcn_sinf(x0);
    float pr, pi;
    gs_sincos2pi(xs, &pr, &pi);
    pi = -pi;
#pragma unroll
    for (int m = 0; m < GS_CSI_NC / 4; ++m) {
        const float vr = fmaf(ur, pr, -(ui * pi)), vi = fmaf(ur, pi, ui * pr);
        put(m, make_float4(ur, ui, vr, vi));
        ur = fmaf(vr, pr, -(vi * pi));
        ui = fmaf(vr, pi, vi * pr);
    }
}

// csi_phasor_half for n path lengths: out[i][k] = (Re, Im) u_k, for its
// accuracy check; one thread per (length, half)
__global__ void __launch_bounds__(256) k_csi_phasors(size_t n, const float* plen, GsCsi t, float4* out)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < 2 * n; i += (size_t)gridDim.x * 256) {
        const size_t e = i >> 1;
        const int h = (int)(i & 1);
        float4* o = out + e * (GS_CSI_NC / 2) + h * (GS_CSI_NC / 4);
        csi_phasor_half(plen[e], h, t, [&](int m, float4 v) { o[m] = v; });
    }
}

hipError_t launch_csi_phasors(size_t n, const float* plen, const GsCsi& t, float* out, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_csi_phasors, dim3((unsigned)std::min<size_t>((2 * n + 255) / 256, 65536)), dim3(256), 0,
                       st, n, plen, t, reinterpret_cast<float4*>(out));
    return hipGetLastError();
}

template <int PPL>
__global__ void __launch_bounds__(GS_WAVE) k_render_fwd_csi(Dims d, RenderCsiArgs a)
{
    __shared__ float4 s0[CSI_BATCH];
    __shared__ float4 s1[CSI_BATCH];
    __shared__ float4 se[CSI_BATCH][CSI_NC2];
    const Pixels<PPL> pm = map_pixels<PPL>(d, a.order, (int)a.total[1]);
    if (!pm.valid) return;
    if ((uint64_t)a.total[0] > a.Kcap) {   // binning overflow, reported by the host
        overflow_outputs<PPL>(d, pm, a.out, nullptr, 2 * GS_CSI_NC, a.finalT, a.ncontrib);
        return;
    }
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
                    csi_phase(gs_cycle_frac(L, a.csi.khi[2 * k2], a.csi.klo[2 * k2]), &c0, &n0);
                    csi_phase(gs_cycle_frac(L, a.csi.khi[2 * k2 + 1], a.csi.klo[2 * k2 + 1]), &c1, &n1);
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
                // power >= thr: alpha >= 1/255 (gs_alpha_thr)
                if (power > 0.0f || power < B.z) continue;
                const float alpha = fminf(0.99f, B.y * gs_expf_blend(power));
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

// Forward on the matrix cores.  Per batch of CSI_MFB staged entries the
// kept ones (block test passed) are compacted into slots 0..nk-1 in list
// order; all 64 lanes then evaluate the kept entries' emissions
// (slot, subcarrier pair) into LDS, and each lane walks the slots for its
// pixel exactly as the VALU forward does (same alpha, T, termination and
// contributor count), writing its blend weight w = alpha T (0 when the pair
// does not contribute) to W[pixel][slot].  The batch's contribution to the
// 64 x 104 map block is then the product W[64 x nk] E[nk x 104], done as
// v_mfma_f32_32x32x2f32 steps (2 pixel blocks x 4 column blocks of 32, the
// last padded from 104 to 128 with zeros) over ceil(nk / 2) entry pairs.  Only the order
// of the sums over entries differs from the VALU forward.
//
// MFMA 32x32x2 f32 operand layout (wave64): lane l supplies A[l%32][l/32] and
// B[l/32][l%32]; accumulator element v of lane l is C[8(v/4) + 4(l/32) +
// v%4][l%32].
// entries staged per batch: 16 keeps the workgroup's LDS (emissions, blend
// weights, slot tables) near 12 KB, so that the registers and not the LDS
// bound the waves per CU (32 entries: 23.5 KB, 1.625 against 1.57 ms at
// 50k x 32)
#define CSI_MFB 16
#define CSI_EW 108            // LDS row of an entry's 104 emissions (+ 4)
#define CSI_WW (CSI_MFB + 1)

typedef float floatx16_csi __attribute__((ext_vector_type(16)));

__global__ void __launch_bounds__(GS_WAVE) __attribute__((amdgpu_waves_per_eu(2)))
k_render_fwd_csi_mfma(Dims d, RenderCsiArgs a)
{
    __shared__ float4 sk0[CSI_MFB];            // kept entries' splat records, by slot
    __shared__ float4 sk1[CSI_MFB];
    __shared__ float2 ssig[CSI_MFB];
    __shared__ float splen[CSI_MFB];
    __shared__ int sjj[CSI_MFB];               // staged position of slot k
    __shared__ __attribute__((aligned(16))) float sE[CSI_MFB][CSI_EW];
    __shared__ float sW[GS_WAVE][CSI_WW];
    const Pixels<1> pm = map_pixels<1>(d, a.order, (int)a.total[1]);
    if (!pm.valid) return;
    if ((uint64_t)a.total[0] > a.Kcap) {   // binning overflow, reported by the host
        overflow_outputs<1>(d, pm, a.out, nullptr, 2 * GS_CSI_NC, a.finalT, a.ncontrib);
        return;
    }
    if (ballot(pm.inside[0]) == 0) return;
    const int r = pm.r;
    const int lane = threadIdx.x;
    const uint2 range = a.ranges[(size_t)r * d.T + pm.tile];
    const uint32_t rb = __builtin_amdgcn_readfirstlane(range.x);
    const uint32_t re = __builtin_amdgcn_readfirstlane(range.y);
    const size_t rP = rx_map(d, r).roff;

    float T = pm.inside[0] ? 1.0f : -1.0f;   // T < 0: finished pixel
    uint32_t last = 0;
    const float pxf = (float)pm.px[0], pyf = (float)pm.py[0];
    floatx16_csi acc[2][4];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nb = 0; nb < 4; ++nb)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[mb][nb][v] = 0.0f;

    for (uint32_t start = rb; start < re; start += CSI_MFB) {
        if (ballot(T > 0.0f) == 0) break;
        // stage and test CSI_MFB entries; compact the kept ones into slots
        const uint32_t e = start + lane;
        bool keep = false;
        float4 A, B;
        size_t gi = 0;
        if (lane < CSI_MFB && e < re) {
            gi = rP + a.plist[e];
            A = a.splat[2 * gi];
            B = a.splat[2 * gi + 1];
            keep = !block_misses(A.x, A.y, A.z, A.w, B.x, B.z, pm.x0, pm.x1, pm.y0, pm.y1);
        }
        const uint64_t mask = ballot(keep);
        const int nk = __popcll(mask);
        __syncthreads();   // the previous batch is done with the slot tables
        if (keep) {
            const int k = __popcll(mask & lanemask_lt());
            sk0[k] = A;
            sk1[k] = B;
            ssig[k] = a.sig[gi];
            splen[k] = a.plen[gi];
            sjj[k] = lane;
        }
        if (nk == 0) continue;
        __syncthreads();
        // emissions of the kept entries, (slot, subcarrier pair) over all lanes;
        // an odd last slot is paired with a zero row
        for (int it = lane; it < nk * CSI_NC2; it += GS_WAVE) {
            const int k = it / CSI_NC2, k2 = it % CSI_NC2;
            const float2 S = ssig[k];
            const float L = splen[k];
            float c0, n0, c1, n1, e0r, e0i, e1r, e1i;
            csi_phase(gs_cycle_frac(L, a.csi.khi[2 * k2], a.csi.klo[2 * k2]), &c0, &n0);
            csi_phase(gs_cycle_frac(L, a.csi.khi[2 * k2 + 1], a.csi.klo[2 * k2 + 1]), &c1, &n1);
            gs_csi_emit(S.x, S.y, c0, n0, &e0r, &e0i);
            gs_csi_emit(S.x, S.y, c1, n1, &e1r, &e1i);
            *reinterpret_cast<float4*>(&sE[k][4 * k2]) = make_float4(e0r, e0i, e1r, e1i);
        }
        if (nk & 1)
            for (int c = lane; c < 2 * GS_CSI_NC; c += GS_WAVE) sE[nk][c] = 0.0f;
        // the lane's pixel against the kept entries in list order
        const uint32_t base = start - rb + 1;
        for (int k = 0; k < nk; ++k) {
            const float4 Ak = sk0[k], Bk = sk1[k];
            float w = 0.0f;
            const float dx = Ak.x - pxf, dy = Ak.y - pyf;
            const float power = pair_power(Ak.z, Ak.w, Bk.x, dx, dy);
            // power >= thr: alpha >= 1/255 (gs_alpha_thr)
            if (!(power > 0.0f || power < Bk.z)) {
                const float alpha = fminf(0.99f, Bk.y * gs_expf_blend(power));
                const float test_T = T * (1.0f - alpha);
                if (test_T < 0.0001f) {
                    T = -fabsf(T);
                } else {
                    w = alpha * T;
                    T = test_T;
                    last = base + (uint32_t)sjj[k];
                }
            }
            sW[lane][k] = w;
        }
        if (nk & 1) sW[lane][nk] = 0.0f;
        __syncthreads();
        // C[64 x 128] += W[64 x nk] E[nk x 128]
        const int half = lane >> 5, l32 = lane & 31;
        for (int k0 = 0; k0 < nk; k0 += 2) {
            const float a0 = sW[l32][k0 + half];
            const float a1 = sW[32 + l32][k0 + half];
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                // columns past the 104 emissions multiply zeros
                const float bv = (nb < 3 || l32 < 2 * GS_CSI_NC - 96) ? sE[k0 + half][32 * nb + l32] : 0.0f;
                acc[0][nb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, bv, acc[0][nb], 0, 0, 0);
                acc[1][nb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, bv, acc[1][nb], 0, 0, 0);
            }
        }
    }
    // the map block: row m of the accumulators is the pixel of lane m
    const int bx = pm.px[0] - (lane & 7), by = pm.py[0] - (lane >> 3);
    const size_t plane = (size_t)r * d.W * d.H;
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nb = 0; nb < 4; ++nb)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int m = 32 * mb + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                const int c = 32 * nb + (lane & 31);
                const int x = bx + (m & 7), y = by + (m >> 3);
                if (c < 2 * GS_CSI_NC && x < d.W && y < d.H)
                    a.out[(plane + (size_t)y * d.W + x) * (2 * GS_CSI_NC) + c] = acc[mb][nb][v];
            }
    if (pm.inside[0]) {
        const size_t p = plane + (size_t)pm.py[0] * d.W + pm.px[0];
        a.finalT[p] = fabsf(T);
        a.ncontrib[p] = last;
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
    static_assert(CSI_BATCH == GS_WAVE / 2, "the two halves of the wave stage the same entries");
    for (int top = maxlast - 1; top >= 0; top -= CSI_BATCH) {
        // entry j = lane % 32 of this batch sits at list position top - j;
        // both halves of the wave load and test it, and half h evaluates the
        // phasors of subcarriers 26h .. 26h + 25 (they weight gradients only:
        // csi_phasor_half's recurrence from the exact phase of the half's
        // first subcarrier)
        const int jl = lane & (CSI_BATCH - 1), hh = lane / CSI_BATCH;
        const int rel = top - jl;
        bool keep = false;
        if (rel >= 0) {
            const uint32_t id = a.plist[rb + rel];
            const size_t gi = rP + id;
            const float4 A = a.splat[2 * gi];
            const float4 B = a.splat[2 * gi + 1];
            if (hh == 0) {
                sid[jl] = id;
                s0[jl] = A;
                s1[jl] = B;
                ss[jl] = a.sig[gi];
            }
            keep = !block_misses(A.x, A.y, A.z, A.w, B.x, B.z, pm.x0, pm.x1, pm.y0, pm.y1);
            if (keep) {
                float4* dst = &su[jl][hh * (CSI_NC2 / 2)];
                csi_phasor_half(a.plen[gi], hh, a.csi, [&](int m, float4 v) { dst[m] = v; });
            }
        }
        uint64_t mask = ballot(keep) & 0xffffffffull;   // the lower half's bits: one per entry
        __syncthreads();
        while (mask) {
            const int j = __builtin_ctzll(mask);
            mask &= mask - 1;
            const float4 A = s0[j];
            const float4 B = s1[j];
            const float2 S = ss[j];
            float v[8] = {-0.0f, -0.0f, -0.0f, -0.0f, -0.0f, -0.0f, -0.0f, -0.0f};
            float vL = -0.0f;
            const bool active = (top - j) < last;
            const float dx = A.x - pxf, dy = A.y - pyf;
            const float power = pair_power(A.z, A.w, B.x, dx, dy);
            // the forward's decision: power in [thr, 0] <=> alpha >= 1/255
            const bool c = active && !(power > 0.0f) && !(power < B.z);
            float G = 0.0f, oG = 0.0f, alpha = 0.0f;
            if (c) {
                G = gs_expf_grad(power);
                oG = B.y * G;
                alpha = fminf(0.99f, oG);
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
                    // the raw sums of the g8 record (the preprocess backward
                    // applies the conic and the factors -1/2, -1, -1/2)
                    v[0] = dG * gdx;
                    v[1] = dG * gdy;
                    v[2] = gdx * dx * dG;
                    v[3] = gdx * dy * dG;
                    v[4] = gdy * dy * dG;
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
    if (ppl == 0)
        hipLaunchKernelGGL(k_render_fwd_csi_mfma, dim3((unsigned)(order_length(d) * 4)), dim3(GS_WAVE), 0, s, d, a);
    else if (ppl == 2)
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
