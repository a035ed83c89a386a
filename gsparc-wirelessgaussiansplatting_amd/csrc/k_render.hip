// Front-to-back alpha blend (renderCUDA, forward.cu:151-249) and its adjoint
// (renderCUDA, backward.cu:442-616) for R receivers.
//
// Work unit = one wave64 workgroup owning PPL of the four 8x8 quarters of a
// 16x16 binning tile; every lane carries PPL independent pixel chains, one per
// quarter.  PPL = 1 gives many small work items and the finest culling block
// (best for one receiver); PPL = 2 and 4 share the staging and culling test of
// each list entry and, in the backward, the 8-value wave reduction among 128
// or 256 pixels (PPL = 2 is best for receiver batches).  The 16x16 binning grid itself is part of the numerics
// and never changes.  A wave stages 64 list entries at a time in LDS, testing
// each against its own pixel block on the way in (block_misses): only entries
// that can reach alpha >= 1/255 somewhere in the block are walked, so the
// per-pixel loop sees a fraction of the tile list.  A wave exits as soon as
// all of its pixels are saturated.  Work-group ids
// are mapped so that all waves of a tile run on the same XCD (shared L2).
//
// Per pair the decision alpha >= 1/255 (forward.cu:219-221) is the comparison
// power >= thr, thr = gs_alpha_thr(opacity) from the preprocess: exactly the
// powers whose o * gs_expf_blend(power) reaches 1/255.  The forward evaluates
// the exponential only there; the backward takes the same comparison, so its
// contributors are the forward's whatever its faster exponential rounds to
// (backward.cu:560-563 re-tests alpha with the forward's own exp).
#include "gs_blend.h"
#include "gs_cull.h"
#include "gs_device.h"
#include "gs_kernels.h"

namespace gs {

template <int PPL, int BD, bool COUNT>
__global__ void __launch_bounds__(GS_WAVE) k_render_fwd(Dims d, RenderFwdArgs a)
{
    // BD list entries are staged per batch: NS per lane (BD = 64, 128), or
    // one on the first BD lanes (BD = 32)
    constexpr int NS = BD > GS_WAVE ? BD / GS_WAVE : 1;
    __shared__ float4 s0[NS * GS_WAVE];
    __shared__ float4 s1[NS * GS_WAVE];
    __shared__ float2 ss[NS * GS_WAVE];
    const Pixels<PPL> pm = map_pixels<PPL>(d, a.order, (int)a.total[1]);
    if (!pm.valid) return;
    if ((uint64_t)a.total[0] > a.Kcap) {   // binning overflow, reported by the host
        overflow_outputs<PPL>(d, pm, a.out_re, a.out_im, 1, a.finalT, a.ncontrib);
        return;
    }
    bool any_inside = false;
#pragma unroll
    for (int k = 0; k < PPL; ++k) any_inside |= pm.inside[k];
    if (ballot(any_inside) == 0) return;
    const int r = pm.r;
    const int lane = threadIdx.x;
    const uint2 range = a.ranges[(size_t)r * d.T + pm.tile];
    const uint32_t rb = __builtin_amdgcn_readfirstlane(range.x);
    const uint32_t re = __builtin_amdgcn_readfirstlane(range.y);
    const size_t rP = rx_map(d, r).roff;

    // T carries the termination flag in its sign: a pixel is done once
    // T < 0, with |T| its final transmittance.  A done pixel can still pass
    // the power and alpha tests, but then T (1 - alpha) < 0 < 1e-4 sends it
    // down the termination path, which leaves T negative: no state changes.
    float T[PPL], are[PPL], aim[PPL], pxf[PPL], pyf[PPL];
    uint32_t last[PPL], stop[PPL], cnt[PPL];
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        T[k] = pm.inside[k] ? 1.0f : -1.0f;
        are[k] = 0.0f; aim[k] = 0.0f;
        last[k] = 0; cnt[k] = 0;
        stop[k] = re - rb;   // entries the reference visits: up to the terminating one, else all
        pxf[k] = (float)pm.px[k];
        pyf[k] = (float)pm.py[k];
    }
    for (uint32_t start = rb; start < re; start += BD) {
        bool live = false;
#pragma unroll
        for (int k = 0; k < PPL; ++k) live |= T[k] > 0.0f;
        if (ballot(live) == 0) break;
        // every load of the batch is issued before the first block test
        bool in[NS];
        size_t gi[NS];
#pragma unroll
        for (int b = 0; b < NS; ++b) {
            const uint32_t e = start + b * GS_WAVE + lane;
            in[b] = e < re && (BD >= GS_WAVE || lane < BD);
            gi[b] = in[b] ? rP + a.plist[e] : 0;
        }
        uint64_t masks[NS];
#pragma unroll
        for (int b = 0; b < NS; ++b) {
            bool keep = false;
            if (in[b]) {
                const float4 A = a.splat[2 * gi[b]];
                const float4 B = a.splat[2 * gi[b] + 1];
                s0[b * GS_WAVE + lane] = A;
                s1[b * GS_WAVE + lane] = B;
                ss[b * GS_WAVE + lane] = a.sig[gi[b]];
                keep = !block_misses(A.x, A.y, A.z, A.w, B.x, B.z, pm.x0, pm.x1, pm.y0, pm.y1);
            }
            masks[b] = ballot(keep);
        }
        __syncthreads();
#pragma unroll
        for (int b = 0; b < NS; ++b) {
            uint64_t mask = masks[b];
            // list position (1-based) of entry j of this sub-batch is base + j
            const uint32_t base = start - rb + 1 + b * GS_WAVE;
            while (mask) {
                const int j = __builtin_ctzll(mask);
                mask &= mask - 1;
                const float4 A = s0[b * GS_WAVE + j];
                const float4 B = s1[b * GS_WAVE + j];
                const float2 S = ss[b * GS_WAVE + j];
#pragma unroll
                for (int k = 0; k < PPL; ++k) {
                    const float dx = A.x - pxf[k], dy = A.y - pyf[k];
                    const float power = pair_power(A.z, A.w, B.x, dx, dy);
                    // power >= thr: alpha >= 1/255 (gs_alpha_thr)
                    if (power > 0.0f || power < B.z) continue;
                    const float alpha = fminf(0.99f, B.y * gs_expf_blend(power));
                    const float test_T = T[k] * (1.0f - alpha);
                    if (test_T < 0.0001f) {
                        if (COUNT && T[k] > 0.0f) stop[k] = base + j;
                        T[k] = -fabsf(T[k]);
                        continue;
                    }
                    const float w = alpha * T[k];
                    are[k] = fmaf(S.x, w, are[k]);
                    aim[k] = fmaf(S.y, w, aim[k]);
                    T[k] = test_T;
                    last[k] = base + j;
                    if (COUNT) ++cnt[k];
                }
            }
        }
        __syncthreads();
    }
    if (COUNT) {
        unsigned long long c = 0, v = 0;
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
            c += cnt[k];
            v += pm.inside[k] ? stop[k] : 0u;
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            c += __shfl_xor(c, off);
            v += __shfl_xor(v, off);
        }
        if (lane == 0) {
            atomicAdd(a.counts + 0, v);
            atomicAdd(a.counts + 2, c);
        }
    }
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        if (!pm.inside[k]) continue;
        const size_t p = (size_t)r * d.W * d.H + (size_t)pm.py[k] * d.W + pm.px[k];
        a.out_re[p] = are[k];
        a.out_im[p] = aim[k];
        a.finalT[p] = fabsf(T[k]);
        a.ncontrib[p] = last[k];
    }
}

// One staged list entry of the backward: the splat record (with the
// Gaussian index in the second record's spare slot) and the emission, 48
// bytes, so that one address serves all three reads.
struct __align__(16) BwdRec {
    float4 A;    // px, py, ha, nb
    float4 B;    // hc, opacity, thr, Gaussian index (bits)
    float2 S;    // emission (real, imag)
    float2 pad;
};

// Per-lane sums over the lane's pixels for the current entry, with
// D = o G dL/dalpha: sh = sum D dx, sk = sum D dy, sxx = sum D dx^2,
// sxy = sum D dx dy, syy = sum D dy^2, sg = sum G dL/dalpha, swr / swi =
// sum w dL/dpix.  The 2D-mean and conic gradients are linear in them.
struct BwdSums {
    float sh, sk, sxx, sxy, syy, sg, swr, swi;
};

// One (entry, quarter) step of the reverse walk (backward.cu:555-610).
// c marks the lanes whose pixel the entry contributes to.
//
// FIRST (the wave's first quarter): every lane runs the arithmetic and the
// step writes the sums instead of adding to them, so that they need no
// clearing between entries.  The lanes without a contribution take G = 0,
// hence alpha = 0: their accumulators are unchanged (fma(0, e, acc) = acc),
// their T is kept by a select, and every term they write is 0.
// Otherwise only the contributing lanes run it (EXEC-masked) and add to the
// sums, which hold the first step's terms or were cleared before.
template <int MODE, bool FIRST>
__device__ __forceinline__ void bwd_step(bool c, float power, float dx, float dy, float opac, float2 S, float& T,
                                         float& acc_re, float& acc_im, float gre, float gim, BwdSums& s)
{
    if (!FIRST && !c) return;
    const float G = (FIRST && !c) ? 0.0f : gs_expf_grad(power);
    const float oG = opac * G;
    const float alpha = fminf(0.99f, oG);
    // T before this Gaussian: T_after / (1 - alpha) (backward.cu:565)
    const float Tb = T * __builtin_amdgcn_rcpf(1.0f - alpha);
    T = (FIRST && !c) ? T : Tb;
    const float sr = S.x;
    const float si = (MODE == GRAD_REFERENCE) ? S.x : S.y;
    // dL/dalpha = (c - accum) dL/dpixel T (backward.cu:586, 592), then
    // accum = alpha c + (1 - alpha) accum for the next contributor
    const float er = sr - acc_re;
    float dalpha = er * gre;
    acc_re = fmaf(alpha, er, acc_re);
    if (MODE == GRAD_EXACT) {
        const float ei = si - acc_im;
        dalpha = fmaf(ei, gim, dalpha);
        acc_im = fmaf(alpha, ei, acc_im);
    }
    dalpha *= T;
    const float w = alpha * T;
    // the exact adjoint has no gradient through the clamped alpha
    const float da = (MODE == GRAD_EXACT && oG > 0.99f) ? 0.0f : dalpha;
    const float D = oG * da;
    const float h = D * dx, q = D * dy;
    if (FIRST) {
        s.swr = w * gre;
        s.swi = w * gim;
        s.sh = h;
        s.sk = q;
        s.sxx = h * dx;
        s.sxy = h * dy;
        s.syy = q * dy;
        s.sg = G * da;
    } else {
        s.swr = fmaf(w, gre, s.swr);
        s.swi = fmaf(w, gim, s.swi);
        s.sh += h;
        s.sk += q;
        s.sxx = fmaf(h, dx, s.sxx);
        s.sxy = fmaf(h, dy, s.sxy);
        s.syy = fmaf(q, dy, s.syy);
        s.sg = fmaf(G, da, s.sg);
    }
}

// DET: the deterministic form (PPL = 4, so that one wave walks each tile
// list): each entry's sums are stored to the entry's instance slot instead
// of added to the record's sums.  A record's slots are a run of its own
// (a.det_off, the exclusive scan of the records' tile counts), its tiles in
// row order; k_det_gather adds each run in that order.  Slots of entries
// without a contribution keep the zeros the host filled them with.
template <int MODE, int PPL, int BD, bool COUNT, bool DET = false>
__global__ void __launch_bounds__(GS_WAVE) k_render_bwd(Dims d, RenderBwdArgs a)
{
    static_assert(!DET || PPL == 4, "the deterministic backward walks each tile with one wave");
    // BD list entries are staged per batch: NS per lane (BD = 64, 128), or
    // one on the first BD lanes (BD = 32)
    constexpr int NS = BD > GS_WAVE ? BD / GS_WAVE : 1;
    __shared__ BwdRec rec[NS * GS_WAVE];
    const Pixels<PPL> pm = map_pixels<PPL>(d, a.order, (int)a.total[1]);
    if (!pm.valid) return;
    if ((uint64_t)a.total[0] > a.Kcap) return;   // no lists: the gradient stays cleared
    const int r = pm.r;
    const int lane = threadIdx.x;
    float T[PPL], gre[PPL], gim[PPL], pxf[PPL], pyf[PPL];
    int last[PPL];
    int maxlast = 0;
    unsigned long long walked = 0, contrib = 0;
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const size_t p = (size_t)r * d.W * d.H + (size_t)pm.py[k] * d.W + pm.px[k];
        T[k] = pm.inside[k] ? a.finalT[p] : 0.0f;
        last[k] = pm.inside[k] ? (int)a.ncontrib[p] : 0;
        gre[k] = pm.inside[k] ? a.dout_re[p] : 0.0f;
        gim[k] = pm.inside[k] ? a.dout_im[p] : 0.0f;
        pxf[k] = (float)pm.px[k];
        pyf[k] = (float)pm.py[k];
        if (COUNT) walked += (unsigned long long)last[k];
    }
    // per quarter, the largest contributor count of its pixels: entries at or
    // past it are walked by none of them
    int qlast[PPL];
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        int m = last[k];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) m = max(m, __shfl_xor(m, off));
        qlast[k] = __builtin_amdgcn_readfirstlane(m);
        maxlast = max(maxlast, qlast[k]);
    }
    const uint2 range = a.ranges[(size_t)r * d.T + pm.tile];
    const uint32_t rb = __builtin_amdgcn_readfirstlane(range.x);
    const size_t rP = rx_map(d, r).roff;

    // accum of backward.cu:579-580, updated right after each contributor
    // instead of before the next one (the same values, no last_* state)
    float acc_re[PPL], acc_im[PPL];
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        acc_re[k] = 0.0f; acc_im[k] = 0.0f;
    }
    float* const g8r = a.g8 + rP * 8;
    // pixel row groups of the wave: quarters k and k ^ 1 share a row when PPL >= 2
    constexpr int NROW = PPL == 1 ? 1 : PPL / 2;
    for (int top = maxlast - 1; top >= 0; top -= BD) {
        // entry j of sub-batch b sits at list position top - 64 b - j (walked
        // back to front); every load of the batch is issued before the tests
        int rel[NS];
        uint32_t id[NS];
#pragma unroll
        for (int b = 0; b < NS; ++b) {
            rel[b] = (BD >= GS_WAVE || lane < BD) ? top - b * GS_WAVE - lane : -1;
            id[b] = rel[b] >= 0 ? a.plist[rb + rel[b]] : 0u;
        }
        uint64_t qmask[NS][PPL];
#pragma unroll
        for (int b = 0; b < NS; ++b) {
            bool hit[PPL];
#pragma unroll
            for (int k = 0; k < PPL; ++k) hit[k] = false;
            if (rel[b] >= 0) {
                const size_t gi = rP + id[b];
                const float4 A = a.splat[2 * gi];
                float4 B = a.splat[2 * gi + 1];
                const float2 S = a.sig[gi];
                B.w = __uint_as_float(id[b]);
                rec[b * GS_WAVE + lane].A = A;
                rec[b * GS_WAVE + lane].B = B;
                rec[b * GS_WAVE + lane].S = S;
                if (DET) {
                    // the instance's slot: the record's run start plus this
                    // tile's row-major index inside the record's rect
                    int x0, y0, x1, y1;
                    get_rect(A.x, A.y, a.radii[gi], d.gx, d.gy, x0, y0, x1, y1);
                    const int tx = pm.tile % d.gx, ty = pm.tile / d.gx;
                    const uint32_t slot = a.det_off[gi] + (uint32_t)((ty - y0) * (x1 - x0) + (tx - x0));
                    rec[b * GS_WAVE + lane].pad.x = __uint_as_float(slot);
                }
                // culled per 8x8 quarter: an entry is walked only over the
                // quarters where it can reach alpha >= 1/255 and where some
                // pixel's contributors reach back to it
#pragma unroll
                for (int k = 0; k < PPL; ++k)
                    hit[k] = rel[b] < qlast[k] && !block_misses(A.x, A.y, A.z, A.w, B.x, B.z, pm.qx[k],
                                                                 pm.qx[k] + 7.0f, pm.qy[k], pm.qy[k] + 7.0f);
            }
#pragma unroll
            for (int k = 0; k < PPL; ++k) qmask[b][k] = ballot(hit[k]);
        }
        __syncthreads();
#pragma unroll
        for (int b = 0; b < NS; ++b) {
            uint64_t mask = 0;
#pragma unroll
            for (int k = 0; k < PPL; ++k) mask |= qmask[b][k];
            while (mask) {
                const int j = __builtin_ctzll(mask);
                mask &= mask - 1;
                const float4 A = rec[b * GS_WAVE + j].A;
                const float4 B = rec[b * GS_WAVE + j].B;
                const float2 S = rec[b * GS_WAVE + j].S;
                const int pos = top - b * GS_WAVE - j;
                // quarters k and k ^ 1 share their pixel rows: the row terms of
                // pair_power are computed once per row
                float dyr[NROW], hdd[NROW];
#pragma unroll
                for (int g = 0; g < NROW; ++g) {
                    dyr[g] = A.y - pyf[g * (PPL / NROW)];
                    hdd[g] = (B.x * dyr[g]) * dyr[g];
                }
                // the entry's sums: written by quarter 0's step, or added to
                // from zero by a later quarter's.  Defined on every path, so
                // that no value merges in from the previous entry: with an
                // undefined start the compiler copies the sums before the
                // in-place lane swaps of every reduction (8 moves each)
                BwdSums s{};
                bool have = false;   // wave-uniform: a quarter of this entry had a contributor
#pragma unroll
                for (int k = 0; k < PPL; ++k) {
                    if (!((qmask[b][k] >> j) & 1ull)) continue;   // wave-uniform: the quarter is culled
                    const int g = k / (PPL / NROW);
                    const float dx = A.x - pxf[k];
                    const float power = fmaf(dx, fmaf(A.w, dyr[g], A.z * dx), hdd[g]);   // pair_power
                    // the forward's decision: power in [thr, 0] <=> alpha >= 1/255
                    const bool act = pos < last[k], nonpos = !(power > 0.0f), reach = !(power < B.z);
                    const bool c = act && nonpos && reach;
                    // the masks of single compares combine on the scalar unit
                    if ((ballot(act) & ballot(nonpos) & ballot(reach)) == 0) continue;
                    if (k == 0) {
                        bwd_step<MODE, true>(c, power, dx, dyr[g], B.y, S, T[k], acc_re[k], acc_im[k], gre[k], gim[k], s);
                    } else {
                        bwd_step<MODE, false>(c, power, dx, dyr[g], B.y, S, T[k], acc_re[k], acc_im[k], gre[k], gim[k], s);
                    }
                    have = true;
                    if (COUNT) contrib += c ? 1u : 0u;
                }
                if (!have) continue;
                // the entry's raw sums (g8 record layout): the preprocess backward
                // applies the conic to the first two (dL/dmean2D, backward.cu:600-605)
                // and the factors -1/2, -1, -1/2 to the next three (backward.cu:608-610)
                float v[8];
                v[0] = s.sh;
                v[1] = s.sk;
                v[2] = s.sxx;
                v[3] = s.sxy;
                v[4] = s.syy;
                v[5] = s.sg;
                v[6] = s.swr;
                v[7] = s.swi;
                float tot = wave_sum8(v, lane);
                // completed on all lanes, so that the last DPP step is one add
                asm volatile("" : "+v"(tot));
                if (DET) {
                    // the only wave walking this list: a plain store to the instance slot
                    const uint32_t slot =
                        __builtin_amdgcn_readfirstlane(__float_as_uint(rec[b * GS_WAVE + j].pad.x));
                    float* ip = a.inst8 + (size_t)slot * 8;
                    if ((lane & 7) == 0) ip[lane >> 3] = tot;
                } else {
                    // the record index is wave-uniform: its address arithmetic stays scalar
                    const uint32_t gid = __builtin_amdgcn_readfirstlane(__float_as_uint(B.w));
                    float* rp = g8r + (size_t)gid * 8;
                    if ((lane & 7) == 0) atomicAdd(rp + (lane >> 3), tot);
                }
            }
        }
        __syncthreads();
    }
    if (COUNT) {
        unsigned long long w = walked, c = contrib;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            w += __shfl_xor(w, off);
            c += __shfl_xor(c, off);
        }
        if (lane == 0) {
            atomicAdd(a.counts + 1, w);
            atomicAdd(a.counts + 3, c);
        }
    }
}

// ---------------------------------------------------------------- deterministic sums
// The instance slots of the deterministic backward: record i (visible, with
// the rect getRect gives its centre and radius, auxiliary.h:14-24, as the
// binning placed it) owns the slots [off[i], off[i] + its tile count), its
// tiles in row order; the counts of all records add up to K.  Three
// launches: per 1024 records their counts' sum, one workgroup's exclusive
// scan of those sums, then each block's exclusive scan from its base.
__device__ __forceinline__ uint32_t det_tiles(const Dims& d, const DetArgs& a, size_t rec)
{
    const int radius = a.radii[rec];
    if (radius <= 0) return 0u;
    const float4 A = a.splat[2 * rec];
    int x0, y0, x1, y1;
    get_rect(A.x, A.y, radius, d.gx, d.gy, x0, y0, x1, y1);
    return (uint32_t)((x1 - x0) * (y1 - y0));
}

__device__ __forceinline__ uint32_t det_block_scan(uint32_t v, uint32_t& total)
{
    __shared__ uint32_t s_wave[1024 / GS_WAVE];
    const int wave = threadIdx.x / GS_WAVE, lane = threadIdx.x % GS_WAVE;
    const uint32_t x = wave_incl_sum(v);
    if (lane == GS_WAVE - 1) s_wave[wave] = x;
    __syncthreads();
    uint32_t before = 0, tot = 0;
    for (int w = 0; w < 1024 / GS_WAVE; ++w) {
        const uint32_t t = s_wave[w];
        if (w < wave) before += t;
        tot += t;
    }
    __syncthreads();
    total = tot;
    return before + x - v;
}

__global__ void __launch_bounds__(1024) k_det_count(Dims d, DetArgs a)
{
    const size_t rec = (size_t)blockIdx.x * 1024 + threadIdx.x;
    const uint32_t n = rec < d.Nrec ? det_tiles(d, a, rec) : 0u;
    uint32_t total;
    det_block_scan(n, total);
    if (threadIdx.x == 0) a.bsum[blockIdx.x] = total;
}

__global__ void __launch_bounds__(1024) k_det_scan(uint32_t nb, DetArgs a)
{
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += 1024) {
        const uint32_t b = b0 + threadIdx.x;
        const uint32_t v = b < nb ? a.bsum[b] : 0u;
        uint32_t total;
        const uint32_t ex = det_block_scan(v, total);
        if (b < nb) a.bsum[b] = carry + ex;
        carry += total;
    }
}

__global__ void __launch_bounds__(1024) k_det_offsets(Dims d, DetArgs a)
{
    const size_t rec = (size_t)blockIdx.x * 1024 + threadIdx.x;
    const uint32_t n = rec < d.Nrec ? det_tiles(d, a, rec) : 0u;
    uint32_t total;
    const uint32_t ex = det_block_scan(n, total);
    if (rec < d.Nrec) a.off[rec] = a.bsum[blockIdx.x] + ex;
}

// Each visible record's run of instance slots summed in order into its raw
// sums g8 (the record layout of the atomic form); the result does not depend
// on timing.  Neighbouring threads read neighbouring runs.
__global__ void __launch_bounds__(256) k_det_gather(Dims d, DetArgs a)
{
    const size_t rec = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (rec >= d.Nrec) return;
    // every record is written (zeros without tiles or lists), so the
    // deterministic backward needs no cleared gradient buffer
    float4* gp = reinterpret_cast<float4*>(a.g8 + rec * 8);
    const uint32_t n = (uint64_t)a.total[0] > a.Kcap ? 0u : det_tiles(d, a, rec);
    if (n == 0) {
        gp[0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        gp[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const float4* ip = reinterpret_cast<const float4*>(a.inst8) + (size_t)a.off[rec] * 2;
    float4 s0 = ip[0], s1 = ip[1];
    for (uint32_t k = 1; k < n; ++k) {
        const float4 u = ip[2 * k], v = ip[2 * k + 1];
        s0.x += u.x; s0.y += u.y; s0.z += u.z; s0.w += u.w;
        s1.x += v.x; s1.y += v.y; s1.z += v.z; s1.w += v.w;
    }
    gp[0] = s0;
    gp[1] = s1;
}

size_t det_scratch_words(const Dims& d)
{
    return d.Nrec + (d.Nrec + 1023) / 1024;
}

hipError_t launch_det_offsets(const Dims& d, const DetArgs& a, hipStream_t s)
{
    if (d.Nrec == 0) return hipSuccess;
    const unsigned nb = (unsigned)((d.Nrec + 1023) / 1024);
    hipLaunchKernelGGL(k_det_count, dim3(nb), dim3(1024), 0, s, d, a);
    hipLaunchKernelGGL(k_det_scan, dim3(1), dim3(1024), 0, s, (uint32_t)nb, a);
    hipLaunchKernelGGL(k_det_offsets, dim3(nb), dim3(1024), 0, s, d, a);
    return hipGetLastError();
}

hipError_t launch_det_gather(const Dims& d, const DetArgs& a, hipStream_t s)
{
    if (d.Nrec == 0) return hipSuccess;
    hipLaunchKernelGGL(k_det_gather, dim3((unsigned)((d.Nrec + 255) / 256)), dim3(256), 0, s, d, a);
    return hipGetLastError();
}

static inline dim3 render_grid(const Dims& d, int ppl)
{
    return dim3((unsigned)(order_length(d) * (4 / ppl)));
}

int choose_ppl(const Dims& d, int requested)
{
    if (requested == 1 || requested == 2 || requested == 4) return requested;
    // half a tile per wave once there are enough tiles to fill the chip with
    // two waves per tile; quarter tiles (finest culling, most waves) below
    return ((long long)d.R * d.T >= 2048) ? 2 : 1;
}

int choose_depth(int requested)
{
    return (requested == 32 || requested == 128) ? requested : 64;
}

template <int PPL, bool COUNT>
static void launch_fwd_depth(const Dims& d, const RenderFwdArgs& a, int depth, hipStream_t s)
{
    const dim3 grid = render_grid(d, PPL);
    if (depth == 32) hipLaunchKernelGGL((k_render_fwd<PPL, 32, COUNT>), grid, dim3(GS_WAVE), 0, s, d, a);
    else if (depth == 128) hipLaunchKernelGGL((k_render_fwd<PPL, 128, COUNT>), grid, dim3(GS_WAVE), 0, s, d, a);
    else hipLaunchKernelGGL((k_render_fwd<PPL, 64, COUNT>), grid, dim3(GS_WAVE), 0, s, d, a);
}

template <bool COUNT>
static void launch_fwd_count(const Dims& d, const RenderFwdArgs& a, int ppl, int depth, hipStream_t s)
{
    if (ppl == 4) launch_fwd_depth<4, COUNT>(d, a, depth, s);
    else if (ppl == 2) launch_fwd_depth<2, COUNT>(d, a, depth, s);
    else launch_fwd_depth<1, COUNT>(d, a, depth, s);
}

hipError_t launch_render_fwd(const Dims& d, const RenderFwdArgs& a, int ppl, int depth, hipStream_t s)
{
    if (d.R == 0 || d.T == 0) return hipSuccess;
    depth = choose_depth(depth);
    if (a.counts) launch_fwd_count<true>(d, a, ppl, depth, s);
    else launch_fwd_count<false>(d, a, ppl, depth, s);
    return hipGetLastError();
}

template <int MODE, int PPL, bool COUNT>
static void launch_bwd_depth(const Dims& d, const RenderBwdArgs& a, int depth, hipStream_t s)
{
    const dim3 grid = render_grid(d, PPL);
    if (depth == 32) hipLaunchKernelGGL((k_render_bwd<MODE, PPL, 32, COUNT>), grid, dim3(GS_WAVE), 0, s, d, a);
    else if (depth == 128) hipLaunchKernelGGL((k_render_bwd<MODE, PPL, 128, COUNT>), grid, dim3(GS_WAVE), 0, s, d, a);
    else hipLaunchKernelGGL((k_render_bwd<MODE, PPL, 64, COUNT>), grid, dim3(GS_WAVE), 0, s, d, a);
}

template <int MODE, bool COUNT>
static void launch_bwd_mode(const Dims& d, const RenderBwdArgs& a, int ppl, int depth, hipStream_t s)
{
    if (ppl == 4) launch_bwd_depth<MODE, 4, COUNT>(d, a, depth, s);
    else if (ppl == 2) launch_bwd_depth<MODE, 2, COUNT>(d, a, depth, s);
    else launch_bwd_depth<MODE, 1, COUNT>(d, a, depth, s);
}

hipError_t launch_render_bwd(const Dims& d, const RenderBwdArgs& a, int ppl, int depth, hipStream_t s)
{
    if (d.R == 0 || d.T == 0) return hipSuccess;
    if (a.inst8) {   // deterministic: one wave per tile, batch depth 64, no counters
        const dim3 grid = render_grid(d, 4);
        if (a.mode == GRAD_REFERENCE)
            hipLaunchKernelGGL((k_render_bwd<GRAD_REFERENCE, 4, 64, false, true>), grid, dim3(GS_WAVE), 0, s, d, a);
        else
            hipLaunchKernelGGL((k_render_bwd<GRAD_EXACT, 4, 64, false, true>), grid, dim3(GS_WAVE), 0, s, d, a);
        return hipGetLastError();
    }
    depth = choose_depth(depth);
    const bool count = a.counts != nullptr;
    if (a.mode == GRAD_REFERENCE) {
        if (count) launch_bwd_mode<GRAD_REFERENCE, true>(d, a, ppl, depth, s);
        else launch_bwd_mode<GRAD_REFERENCE, false>(d, a, ppl, depth, s);
    } else {
        if (count) launch_bwd_mode<GRAD_EXACT, true>(d, a, ppl, depth, s);
        else launch_bwd_mode<GRAD_EXACT, false>(d, a, ppl, depth, s);
    }
    return hipGetLastError();
}

}  // namespace gs
