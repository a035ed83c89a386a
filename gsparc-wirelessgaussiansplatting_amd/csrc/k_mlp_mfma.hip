// Shared emission MLP of the single-receiver training configuration on the
// matrix cores (FP32 MFMA, FP32 accumulate).  No counterpart in the reference
// (whose emission MLP is per Gaussian, 3->32->2, gaussian.py:47).
//
// For receiver r and Gaussian g with mean mu:
//   PE(mu) [63]   mu, then sin(pi x), cos(pi x) of x = 2^k mu reduced exactly
//                 to [-1, 1] (x - 2 rint(x / 2)), k = 0..9, each a 3-vector
//   f [5]         d = |rx - mu|, 1 / max(d^2, 0.04), (tx - mu) / |tx - mu|
//   h1 = relu(pe1[g] + W1rx f)      pe1 = PE(mu) W1pe^T + b1, W1 = [W1pe | W1rx]
//   h2 = relu(W2 h1 + b2)           a = w3 . h2 + b3,  emission (a, 0)
// Kernels:
//   k_mlp1_pe       pe1 for 128 Gaussians per workgroup: PE into LDS, then
//                   32x32x2 MFMA tiles against W1pe^T (K = 63, padded to 64)
//   k_mlp1_fwd      a for 128-row tiles of (Gaussian, receiver) rows
//   k_mlp1_bwd      64-row tiles: recompute h1, h2; g2 = da w3 (h2 > 0);
//                   dW2 += g2^T h1, dh1 = (g2 W2)(h1 > 0) as MFMA; from the
//                   dh1 tile in LDS: dW1rx += dh1^T f, the sums over receivers
//                   dpe1[g] += sum_r dh1 and the receiver-feature chain rule
//                   into dmu (atomics once per run of rows of one Gaussian)
//   k_mlp1_pe_bwd   per 128 Gaussians: dPE = dpe1 W1pe and dW1pe += dpe1^T PE
//                   as MFMA, db1 += dpe1, and the PE chain rule into dmu
// Rows are Gaussian-major inside the kernels, q = g R + r, so that a
// Gaussian's receivers sit in consecutive rows; the emission and its
// gradient keep the rasterizer's layout [R, P, 2] (row r P + g).
// Per-workgroup partial sums are reduced by a column-sum kernel in
// workgroup order.
//
// MFMA 32x32x2 f32 operand layout (wave64): lane l supplies A[l%32][l/32] and
// B[l/32][l%32]; accumulator element v of lane l is C[8(v/4) + 4(l/32) +
// v%4][l%32].  LDS rows are padded to an odd stride so that the 32 lanes of a
// half-wave hit 32 different banks in every operand read.
#include <cmath>

#include "gs_blend.h"
#include "gs_device.h"
#include "gs_kernels.h"

namespace gs {

#define M1_H 128     // hidden width
#define M1_S 129     // LDS row stride of 128-wide tiles
#define M1_IN 68     // inputs: 63 PE + 5 receiver features
#define M1_PE 63
#define M1_FEAT 5
#define M1_KS 65     // LDS row stride of 64-wide (PE) tiles
#define M1_CH 9      // chain-rule terms per row (rx_features_ch)

typedef float floatx16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ floatx16 mfma32x32x2(float a, float b, floatx16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int acc_row(int v, int lane) { return 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3); }

// Sums each of p[0..15] over the 32 lanes of each half-wave; lane l returns
// the total of p[(l >> 1) & 15] of its half (lanes l and l ^ 1 both hold it).
// Transpose reduction: the exchanges across 16 lanes (lane swap), 8 lanes
// (row_ror:8), the half-row mirror and quad_perm [2,3,0,1] each halve the
// values a lane carries (8, 4, 2, 1), each pairing lanes that carry the same
// values; quad_perm [1,0,3,2] adds the last pair.
__device__ __forceinline__ float half_sum16(const float (&p)[16], int lane)
{
    const bool b3 = lane & 8, b2 = lane & 4, b1 = lane & 2;
    float w8[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) w8[i] = swap_sum16(p[i], p[i + 8]);   // bit 4 clear: p[i]; set: p[i + 8]
    float x4[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) x4[i] = (b3 ? w8[i + 4] : w8[i]) + dpp_move<0x128>(b3 ? w8[i] : w8[i + 4]);
    float y2[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) y2[j] = (b2 ? x4[j + 2] : x4[j]) + dpp_move<0x141>(b2 ? x4[j] : x4[j + 2]);
    float z = (b1 ? y2[1] : y2[0]) + dpp_move<0x4e>(b1 ? y2[0] : y2[1]);
    return z + dpp_move<0xb1>(z);
}

// acc[t] += sum over NSTEP steps of mfma(A(step), B(step, t)), with the
// operands read in blocks of KB steps: the next block is read (from LDS)
// while this block's products run; the scheduling barriers keep the reads
// ahead of the products.
template <int NACC, int NSTEP, int KB, class FA, class FB>
__device__ __forceinline__ void mfma_steps(floatx16 (&acc)[NACC], FA a_of, FB b_of)
{
    static_assert(NSTEP % KB == 0, "whole blocks");
    constexpr int NB = NSTEP / KB;
    float ab[2][KB], bb[2][KB][NACC];
#pragma unroll
    for (int i = 0; i < KB; ++i) {
        ab[0][i] = a_of(i);
#pragma unroll
        for (int t = 0; t < NACC; ++t) bb[0][i][t] = b_of(i, t);
    }
#pragma unroll
    for (int kb = 0; kb < NB; ++kb) {
        const int cur = kb & 1, nxt = cur ^ 1;
        if (kb + 1 < NB) {
#pragma unroll
            for (int i = 0; i < KB; ++i) {
                const int st = (kb + 1) * KB + i;
                ab[nxt][i] = a_of(st);
#pragma unroll
                for (int t = 0; t < NACC; ++t) bb[nxt][i][t] = b_of(st, t);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < KB; ++i)
#pragma unroll
            for (int t = 0; t < NACC; ++t) acc[t] = mfma32x32x2(ab[cur][i], bb[cur][i][t], acc[t]);
        __builtin_amdgcn_sched_barrier(0);
    }
}

// W2 [n][k] (global, row-major) -> w2t[k * M1_S + n]
__device__ __forceinline__ void load_w2t(const float* w2, float* w2t)
{
    for (int i = threadIdx.x; i < M1_H * M1_H; i += blockDim.x) {
        const int n = i >> 7, k = i & (M1_H - 1);
        w2t[k * M1_S + n] = w2[i];
    }
}

// sin(pi x), cos(pi x) of x = 2^k mu_j reduced exactly to [-1, 1]: the
// hardware v_sin_f32 / v_cos_f32 of x / 2 turns (exact halving; at most
// 1.3e-7 absolute error over the half-turn domain, the routine the
// multi-frequency kernels use and the test suite checks on every float).
__device__ __forceinline__ void pe_sincos(float mu, int k, float* sn, float* cs)
{
    float x = mu * (float)(1 << k);
    x = x - 2.0f * rintf(0.5f * x);
    *sn = __builtin_amdgcn_sinf(0.5f * x);
    *cs = __builtin_amdgcn_cosf(0.5f * x);
}

// PE rows of the 128 Gaussians g0 .. g0+127 into pes[m * M1_KS + c] (c < 64,
// 63 the zero pad; zero rows past P): item (m, b) of 128 x 11 writes block b
// of row m, b = 0 the mean itself, b = 1 + k the sines and cosines of
// frequency k, its three coordinates together.
__device__ __forceinline__ void pe_rows(const Mlp1Args& a, int g0, float* pes)
{
    for (int i = threadIdx.x; i < 128 * 11; i += blockDim.x) {
        const int m = i / 11, b = i % 11, g = g0 + m;
        float* row = pes + m * M1_KS;
        const bool in = g < a.P;
        const float mu[3] = {in ? a.means[3 * (size_t)g] : 0.0f, in ? a.means[3 * (size_t)g + 1] : 0.0f,
                             in ? a.means[3 * (size_t)g + 2] : 0.0f};
        if (b == 0) {
            row[0] = mu[0];
            row[1] = mu[1];
            row[2] = mu[2];
            row[M1_PE] = 0.0f;
        } else {
            const int k = b - 1;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                float sn, cs;
                pe_sincos(mu[j], k, &sn, &cs);
                row[3 + 6 * k + j] = in ? sn : 0.0f;
                row[6 + 6 * k + j] = in ? cs : 0.0f;
            }
        }
    }
}

// Receiver features of row (g, r).
__device__ __forceinline__ void rx_features(const Mlp1Args& a, const float* mu, const float* rxv, float* f)
{
    const float mx = mu[0], my = mu[1], mz = mu[2];
    const float dx = rxv[0] - mx, dy = rxv[1] - my, dz = rxv[2] - mz;
    const float d2 = dx * dx + dy * dy + dz * dz;
    f[0] = sqrtf(d2);
    f[1] = 1.0f / fmaxf(d2, 0.04f);
    const float tx = a.tx[0] - mx, ty = a.tx[1] - my, tz = a.tx[2] - mz;
    const float tn = sqrtf(tx * tx + ty * ty + tz * tz);
    f[2] = tx / tn;
    f[3] = ty / tn;
    f[4] = tz / tn;
}

// The receiver-feature chain rule's per-row terms, from the same
// intermediate values as rx_features: ch = [rx - mu (3), 1/d (0 at d = 0),
// 2/d2^2 (0 where the 1/max(d2, 0.04) feature is clamped), 1/|tx - mu|,
// the unit TX direction (3)].
__device__ __forceinline__ void rx_features_ch(const Mlp1Args& a, const float* mu, const float* rxv, float* f, float* ch)
{
    const float mx = mu[0], my = mu[1], mz = mu[2];
    const float dx = rxv[0] - mx, dy = rxv[1] - my, dz = rxv[2] - mz;
    const float d2 = dx * dx + dy * dy + dz * dz;
    const float d = sqrtf(d2);
    f[0] = d;
    f[1] = 1.0f / fmaxf(d2, 0.04f);
    const float tx = a.tx[0] - mx, ty = a.tx[1] - my, tz = a.tx[2] - mz;
    const float tn = sqrtf(tx * tx + ty * ty + tz * tz);
    f[2] = tx / tn;
    f[3] = ty / tn;
    f[4] = tz / tn;
    ch[0] = dx;
    ch[1] = dy;
    ch[2] = dz;
    ch[3] = d > 0.0f ? 1.0f / d : 0.0f;
    ch[4] = d2 > 0.04f ? 2.0f / (d2 * d2) : 0.0f;
    ch[5] = 1.0f / tn;
    ch[6] = f[2];
    ch[7] = f[3];
    ch[8] = f[4];
}

// ------------------------------------------------------------ PE layer
// pe1[g][n] = sum_c PE(mu_g)[c] W1[n][c] + b1[n], 128 Gaussians per tile,
// wave w the tile's Gaussians 32w..32w+31.
__global__ void __launch_bounds__(256) k_mlp1_pe(Mlp1Args a)
{
    __shared__ float wpt[64 * M1_S];       // W1pe^T: wpt[c * M1_S + n] = W1[n][c], row 63 zero
    __shared__ float pes[128 * M1_KS];     // PE rows
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < 64 * M1_H; i += 256) {
        const int n = i >> 6, c = i & 63;
        wpt[c * M1_S + n] = c < M1_PE ? a.w1[n * M1_IN + c] : 0.0f;
    }
    for (int tile = blockIdx.x; tile * 128 < a.P; tile += gridDim.x) {
        const int g0 = tile * 128;
        __syncthreads();
        pe_rows(a, g0, pes);
        __syncthreads();
        floatx16 acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t] = floatx16{};
        const float* arow = pes + (32 * w + (lane & 31)) * M1_KS + (lane >> 5);
        const float* bcol = wpt + (lane >> 5) * M1_S + (lane & 31);
        for (int ks = 0; ks < 32; ++ks) {
            const float av = arow[2 * ks];
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = mfma32x32x2(av, bcol[2 * ks * M1_S + 32 * t], acc[t]);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int n = 32 * t + (lane & 31);
            const float bb = a.b1[n];
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int g = g0 + 32 * w + acc_row(v, lane);
                if (g < a.P) a.pe1[(size_t)g * M1_H + n] = acc[t][v] + bb;
            }
        }
    }
}

// Layer 1 of ROWS rows starting at q0, in two steps with a barrier between.
// l1_rows: thread t < ROWS computes row t's receiver features into
// fs[t * M1_FEAT + i] and, when chs is given, its chain-rule terms into
// chs[t * M1_CH + i] (zeros past the last row).
// l1_units: thread t takes the 8 units 8 (t >> 4) .. + 8, whose W1rx
// weights it holds in registers for the whole kernel, and the rows
// (t & 15) + 16 j: h1 = relu(pe1 + W1rx f) into h1s[m * M1_S + u].  The
// rows' pe1 slices are read as float4s (pe1 rows are 512-byte aligned),
// all before the first use.
template <int ROWS>
__device__ __forceinline__ void l1_rows(const Mlp1Args& a, long long q0, float* fs, float* chs)
{
    const int m = threadIdx.x;
    if (m >= ROWS) return;
    const long long q = q0 + m;
    float f[M1_FEAT] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float ch[M1_CH] = {};
    if (q < (long long)a.P * a.R) {
        const int g = (int)q / a.R, r = (int)q - g * a.R;
        float mu[3], rxv[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            mu[i] = a.means[3 * (size_t)g + i];
            rxv[i] = a.rx[3 * r + i];
        }
        if (chs) rx_features_ch(a, mu, rxv, f, ch);
        else rx_features(a, mu, rxv, f);
    }
#pragma unroll
    for (int i = 0; i < M1_FEAT; ++i) fs[m * M1_FEAT + i] = f[i];
    if (chs)
#pragma unroll
        for (int i = 0; i < M1_CH; ++i) chs[m * M1_CH + i] = ch[i];
}

// this thread's 8 units' W1rx rows
__device__ __forceinline__ void l1_weights(const Mlp1Args& a, float (&wr)[8][M1_FEAT])
{
    const int u0 = 8 * (threadIdx.x >> 4);
#pragma unroll
    for (int e = 0; e < 8; ++e)
#pragma unroll
        for (int i = 0; i < M1_FEAT; ++i) wr[e][i] = a.w1[(u0 + e) * M1_IN + M1_PE + i];
}

template <int ROWS>
__device__ __forceinline__ void l1_units(const Mlp1Args& a, long long q0, const float (&wr)[8][M1_FEAT],
                                         const float* fs, float* h1s)
{
    constexpr int RPT = ROWS / 16;   // rows per thread
    const int rb = threadIdx.x & 15, u0 = 8 * (threadIdx.x >> 4);
    const long long n_rows = (long long)a.P * a.R;
    float4 pv[RPT][2];
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const long long q = q0 + rb + 16 * j;
        const int g = q < n_rows ? (int)q / a.R : 0;
        const float4* pe4 = reinterpret_cast<const float4*>(a.pe1 + (size_t)g * M1_H + u0);
        pv[j][0] = pe4[0];
        pv[j][1] = pe4[1];
    }
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int m = rb + 16 * j;
        const bool valid = q0 + m < n_rows;
        float f[M1_FEAT];
#pragma unroll
        for (int i = 0; i < M1_FEAT; ++i) f[i] = fs[m * M1_FEAT + i];
        const float pj[8] = {pv[j][0].x, pv[j][0].y, pv[j][0].z, pv[j][0].w,
                             pv[j][1].x, pv[j][1].y, pv[j][1].z, pv[j][1].w};
        float* dst = h1s + m * M1_S + u0;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            float v = pj[e];
#pragma unroll
            for (int i = 0; i < M1_FEAT; ++i) v = fmaf(wr[e][i], f[i], v);
            dst[e] = valid ? fmaxf(v, 0.0f) : 0.0f;
        }
    }
}

// The backward's layer 1 (64-row tiles): thread t takes row t % 64 and the
// 32 units 32 (t / 64) .. + 32 (W1rx from LDS), computing the row's features
// itself, so that no thread waits for another's.  layer1_load reads the
// row's pe1 slice (float4s), mean and receiver, a tile ahead; layer1_store
// writes h1 into h1s[m * M1_S + u], the features into fs[m * M1_FEAT + i]
// and the chain-rule terms into chs[m * M1_CH + i].
struct L1In {
    bool valid;
    float4 pv[8];
    float mu[3], rxv[3];
    float da;   // dL/da of the row (the rasterizer's emission gradient, real part)
};

__device__ __forceinline__ void layer1_load(const Mlp1Args& a, const float* dem, long long q0, L1In& in)
{
    const int m = threadIdx.x & 63, part = threadIdx.x >> 6;
    const long long q = q0 + m;
    in.valid = q < (long long)a.P * a.R;
    const int qq = in.valid ? (int)q : 0;
    const int g = qq / a.R, r = qq - g * a.R;
    const float4* pe4 = reinterpret_cast<const float4*>(a.pe1 + (size_t)g * M1_H + 32 * part);
#pragma unroll
    for (int j = 0; j < 8; ++j) in.pv[j] = pe4[j];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        in.mu[i] = a.means[3 * (size_t)g + i];
        in.rxv[i] = a.rx[3 * r + i];
    }
    in.da = in.valid ? dem[2 * ((size_t)r * a.P + g)] : 0.0f;
}

__device__ __forceinline__ void layer1_store(const Mlp1Args& a, const L1In& in, const float* w1rx_s, float* h1s,
                                             float* fs, float* chs, float* das)
{
    const int m = threadIdx.x & 63, part = threadIdx.x >> 6;
    const int u0 = 32 * part;
    float* dst = h1s + m * M1_S;
    if (part == 0) das[m] = in.da;
    if (!in.valid) {
#pragma unroll
        for (int u = 0; u < 32; ++u) dst[u0 + u] = 0.0f;
        if (part == 0)
            for (int i = 0; i < M1_FEAT; ++i) fs[m * M1_FEAT + i] = 0.0f;
        return;
    }
    float f[M1_FEAT], ch[M1_CH];
    rx_features_ch(a, in.mu, in.rxv, f, ch);
    if (part == 0) {
        for (int i = 0; i < M1_CH; ++i) chs[m * M1_CH + i] = ch[i];
        for (int i = 0; i < M1_FEAT; ++i) fs[m * M1_FEAT + i] = f[i];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const float pj[4] = {in.pv[j].x, in.pv[j].y, in.pv[j].z, in.pv[j].w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int u = u0 + 4 * j + e;
            const float* wr = w1rx_s + u * M1_FEAT;
            float v = pj[e];
#pragma unroll
            for (int i = 0; i < M1_FEAT; ++i) v = fmaf(wr[i], f[i], v);
            dst[u] = fmaxf(v, 0.0f);
        }
    }
}

__device__ __forceinline__ void load_w1rx(const float* w1, float* w1rx_s)
{
    for (int i = threadIdx.x; i < M1_H * M1_FEAT; i += blockDim.x) {
        const int u = i / M1_FEAT, k = i % M1_FEAT;
        w1rx_s[i] = w1[u * M1_IN + M1_PE + k];
    }
}

__global__ void __launch_bounds__(256) k_mlp1_fwd(Mlp1Args a)
{
    __shared__ float w2t[M1_H * M1_S];
    __shared__ float h1s[128 * M1_S];
    __shared__ float fs[128 * M1_FEAT];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long long n_rows = (long long)a.P * a.R;
    load_w2t(a.w2, w2t);   // once per workgroup: the grid strides over 128-row tiles
    float wr[8][M1_FEAT];
    l1_weights(a, wr);
    const float b3 = a.b3[0];
    const float* hrow = h1s + (32 * w + (lane & 31)) * M1_S + (lane >> 5);
    const float* bcol = w2t + (lane >> 5) * M1_S + (lane & 31);
    for (long long tile = blockIdx.x; tile * 128 < n_rows; tile += gridDim.x) {
        const long long q0 = tile * 128;
        l1_rows<128>(a, q0, fs, nullptr);
        __syncthreads();   // weights loaded, the features written / the previous tile's h1s reads done
        l1_units<128>(a, q0, wr, fs, h1s);
        __syncthreads();
        floatx16 acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t] = floatx16{};
        mfma_steps<4, M1_H / 2, 4>(
            acc, [&](int ks) { return hrow[2 * ks]; }, [&](int ks, int t) { return bcol[2 * ks * M1_S + 32 * t]; });
        // a = w3 . relu(h2 + b2): each lane's part over its 4 columns, then
        // the sum over the 32 lanes of each half-wave (column l % 32)
        float part[16];
#pragma unroll
        for (int v = 0; v < 16; ++v) part[v] = 0.0f;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int nn = 32 * t + (lane & 31);
            const float bb = a.b2[nn], ww = a.w3[nn];
#pragma unroll
            for (int v = 0; v < 16; ++v) part[v] = fmaf(ww, fmaxf(acc[t][v] + bb, 0.0f), part[v]);
        }
        const float tot = half_sum16(part, lane);
        const int v = (lane >> 1) & 15;
        const int q = (int)q0 + 32 * w + acc_row(v, lane);
        if ((lane & 1) == 0 && q < n_rows) {
            const int g = q / a.R, r = q - g * a.R;
            const size_t o = (size_t)r * a.P + g;
            a.out[2 * o] = tot + b3;
            a.out[2 * o + 1] = 0.0f;
        }
    }
}

// row-kernel partial record per workgroup:
//   dW2 [128*128] | dw3 [128] | db2 [128] | db3 | dW1rx [128*5]
#define M1_PART (M1_H * M1_H + 2 * M1_H + 1 + M1_H * M1_FEAT)

__global__ void __launch_bounds__(256) k_mlp1_bwd(Mlp1Args a, const float* dem, float* dpe1, float* dmu, float* part)
{
    __shared__ float w2t[M1_H * M1_S];
    __shared__ float h1s[64 * M1_S];
    __shared__ float g2s[64 * M1_S];
    __shared__ float w1rx_s[M1_H * M1_FEAT];
    __shared__ float fs[64 * M1_FEAT];
    __shared__ float chs[2][64 * M1_CH];
    __shared__ float das[64];              // dL/da of the tile's rows   // per tile parity: the next tile's layer 1 may run beside this tile's chain rule
    __shared__ float red[4][2][2][32];
    __shared__ float red3[256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int mh = w & 1, nh = w >> 1;     // this wave's row half and pair of 32-column tiles
    const long long n_rows = (long long)a.P * a.R;
    load_w2t(a.w2, w2t);
    load_w1rx(a.w1, w1rx_s);
    floatx16 dw2[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) dw2[u] = floatx16{};
    float dw3p[2] = {0.0f, 0.0f}, db2p[2] = {0.0f, 0.0f}, db3p = 0.0f;
    float dw1rx[M1_FEAT] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    __syncthreads();   // the weights are loaded
    int par = 0;
    L1In in;
    if ((long long)blockIdx.x * 64 < n_rows) layer1_load(a, dem, (long long)blockIdx.x * 64, in);
    for (long long tile = blockIdx.x; tile * 64 < n_rows; tile += gridDim.x, par ^= 1) {
        const long long q0 = tile * 64;
        // No barrier here: every read of h1s and fs of the previous tile
        // came before its last barrier, and its chain rule (wave 0, after
        // that barrier) reads only g2s and chs[par ^ 1], which this layer 1
        // does not write (g2s is next written after this tile's first barrier).
        layer1_store(a, in, w1rx_s, h1s, fs, chs[par], das);
        if (threadIdx.x < 64) db3p += in.da;
        __syncthreads();
        float dav[16];
#pragma unroll
        for (int v = 0; v < 16; ++v) dav[v] = das[32 * mh + acc_row(v, lane)];
        // h2 for rows 32*mh.., columns of tiles 2*nh, 2*nh + 1
        floatx16 c[2] = {floatx16{}, floatx16{}};
        const float* hrow = h1s + (32 * mh + (lane & 31)) * M1_S + (lane >> 5);
        const float* bcol = w2t + (lane >> 5) * M1_S + 64 * nh + (lane & 31);
        mfma_steps<2, M1_H / 2, 8>(
            c, [&](int ks) { return hrow[2 * ks]; }, [&](int ks, int t) { return bcol[2 * ks * M1_S + 32 * t]; });
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int nn = 64 * nh + 32 * t + (lane & 31);
            const float bb = a.b2[nn], ww = a.w3[nn];
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int m = 32 * mh + acc_row(v, lane);
                const float h2 = fmaxf(c[t][v] + bb, 0.0f);
                const float g = h2 > 0.0f ? dav[v] * ww : 0.0f;
                g2s[m * M1_S + nn] = g;
                dw3p[t] = fmaf(dav[v], h2, dw3p[t]);
                db2p[t] += g;
            }
        }
        __syncthreads();
        // dW2[n][k] += sum_m g2[m][n] h1[m][k]: this wave's n in 32*w..+32
        {
            const float* ga = g2s + (lane >> 5) * M1_S + 32 * w + (lane & 31);
            const float* hb = h1s + (lane >> 5) * M1_S + (lane & 31);
            mfma_steps<4, 32, 4>(
                dw2, [&](int ms) { return ga[2 * ms * M1_S]; }, [&](int ms, int u) { return hb[2 * ms * M1_S + 32 * u]; });
        }
        // dh1[m][k] = (sum_n g2[m][n] W2[n][k]) (h1 > 0): rows 32*mh.., column tiles 2*nh, 2*nh + 1
        floatx16 e[2] = {floatx16{}, floatx16{}};
        {
            const float* grow = g2s + (32 * mh + (lane & 31)) * M1_S + (lane >> 5);
            const float* wk = w2t + (64 * nh + (lane & 31)) * M1_S + (lane >> 5);
            mfma_steps<2, M1_H / 2, 8>(
                e, [&](int ns) { return grow[2 * ns]; }, [&](int ns, int t) { return wk[32 * t * M1_S + 2 * ns]; });
        }
        // the next tile's global reads, issued under the rest of this tile
        {
            const long long nt = tile + gridDim.x;
            if (nt * 64 < n_rows) layer1_load(a, dem, nt * 64, in);
        }
        __syncthreads();   // every wave's dW2 reads of h1s are done: dh1 replaces h1 in place
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int k = 64 * nh + 32 * t + (lane & 31);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int m = 32 * mh + acc_row(v, lane);
                float* p = h1s + m * M1_S + k;
                *p = *p > 0.0f ? e[t][v] : 0.0f;
            }
        }
        __syncthreads();
        // from the dh1 tile: dW1rx, the per-Gaussian sums dpe1, and the
        // receiver-feature chain rule.  Row m of the tile is (Gaussian
        // g0 + (r0 + m) / R, receiver (r0 + m) % R), tracked incrementally.
        const int qt = (int)q0, g0 = qt / a.R, r0 = qt - g0 * a.R;
        {
            const int u = threadIdx.x & (M1_H - 1), half = threadIdx.x >> 7;
            const int rs = r0 + 32 * half;
            int g = g0 + rs / a.R, r = rs % a.R;
            float run = 0.0f;
            int gprev = g;
            // the 32 rows' values are read before the (sequential) sums
            float dhv[32];
#pragma unroll
            for (int mm = 0; mm < 32; ++mm) dhv[mm] = h1s[(32 * half + mm) * M1_S + u];
            // the valid rows are a prefix of the tile: past it dh = 0 and no
            // Gaussian's run is flushed (the last one is, after the loop)
#pragma unroll
            for (int mm = 0; mm < 32; ++mm) {
                const int m = 32 * half + mm;
                const bool valid = qt + m < n_rows;
                const float dh = valid ? dhv[mm] : 0.0f;
#pragma unroll
                for (int i = 0; i < M1_FEAT; ++i) dw1rx[i] = fmaf(dh, fs[m * M1_FEAT + i], dw1rx[i]);
                if (valid && g != gprev) {
                    atomicAdd(dpe1 + (size_t)gprev * M1_H + u, run);
                    run = 0.0f;
                    gprev = g;
                }
                run += dh;
                if (++r == a.R) { r = 0; ++g; }
            }
            if (qt + 32 * half < n_rows) atomicAdd(dpe1 + (size_t)gprev * M1_H + u, run);
        }
        {
            // drxf = dh1 W1rx: wave qu sums units 32 qu.. for all 64 rows (lane = row)
            const int m = threadIdx.x & 63, qu = threadIdx.x >> 6;
            float df[M1_FEAT] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 8
            for (int u = 32 * qu; u < 32 * qu + 32; ++u) {
                const float dh = h1s[m * M1_S + u];
#pragma unroll
                for (int i = 0; i < M1_FEAT; ++i) df[i] = fmaf(dh, w1rx_s[u * M1_FEAT + i], df[i]);
            }
#pragma unroll
            for (int i = 0; i < M1_FEAT; ++i) g2s[(qu * M1_FEAT + i) * 64 + m] = df[i];
        }
        __syncthreads();
        if (threadIdx.x < 64) {
            // wave 0, lane = row: the chain rule, then a segmented sum over
            // the lanes of one Gaussian and one atomic per segment
            const int m = threadIdx.x;
            float df[M1_FEAT];
#pragma unroll
            for (int i = 0; i < M1_FEAT; ++i)
                df[i] = g2s[i * 64 + m] + g2s[(M1_FEAT + i) * 64 + m] + g2s[(2 * M1_FEAT + i) * 64 + m] +
                        g2s[(3 * M1_FEAT + i) * 64 + m];
            const int rr = r0 + m;
            const int g = g0 + rr / a.R, r = rr - (rr / a.R) * a.R;
            const bool valid = qt + m < n_rows;
            float gm[3] = {0.0f, 0.0f, 0.0f};
            if (valid) {
                // the row's terms as layer 1 computed them
                const float* ch = chs[par] + m * M1_CH;
                const float dv[3] = {ch[0], ch[1], ch[2]};
                // d(d)/dmu = -diff / d; d(1/d2)/dmu = 2 diff / d2^2 where d2 > 0.04
                const float cd = -df[0] * ch[3];
                const float ci = df[1] * ch[4];
                const float rtn = ch[5];
                const float u3[3] = {ch[6], ch[7], ch[8]};
                const float ug = u3[0] * df[2] + u3[1] * df[3] + u3[2] * df[4];
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    // d(t/|t|)/dmu = -(I - u u^T) / |t|
                    const float gt = -(df[2 + j] - u3[j] * ug) * rtn;
                    gm[j] = (cd + ci) * dv[j] + gt;
                }
            }
            const int gk = valid ? g : -1;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const int gp = __shfl_up(gk, off);
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const float vp = __shfl_up(gm[j], off);
                    if (m >= off && gp == gk) gm[j] += vp;
                }
            }
            const int gn = __shfl_down(gk, 1);
            if (valid && (m == 63 || gn != gk))
#pragma unroll
                for (int j = 0; j < 3; ++j) atomicAdd(dmu + 3 * (size_t)g + j, gm[j]);
        }
    }
    // this workgroup's partial sums
    float* out = part + (size_t)blockIdx.x * M1_PART;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int n = 32 * w + acc_row(v, lane), k = 32 * u + (lane & 31);
            out[n * M1_H + k] = dw2[u][v];
        }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        dw3p[t] += __shfl_xor(dw3p[t], 32);
        db2p[t] += __shfl_xor(db2p[t], 32);
        if (lane < 32) {
            red[w][t][0][lane] = dw3p[t];
            red[w][t][1][lane] = db2p[t];
        }
    }
    red3[threadIdx.x] = db3p;
    __syncthreads();
    if (threadIdx.x < M1_H) {
        // column n = 64*nh + 32*t + j is held by waves (mh = 0, nh) and (mh = 1, nh)
        const int n = threadIdx.x, nh2 = n >> 6, t = (n >> 5) & 1, j = n & 31;
        out[M1_H * M1_H + n] = red[2 * nh2][t][0][j] + red[2 * nh2 + 1][t][0][j];
        out[M1_H * M1_H + M1_H + n] = red[2 * nh2][t][1][j] + red[2 * nh2 + 1][t][1][j];
    }
    if (threadIdx.x == 0) {
        float s = 0.0f;
        for (int i = 0; i < 64; ++i) s += red3[i];
        out[M1_H * M1_H + 2 * M1_H] = s;
    }
    // dW1rx[u][i]: the two row halves of unit u
    __syncthreads();
    float* dws = &red3[0];   // reuse: 256 threads x 5 values do not fit; two rounds of 256
    for (int i = 0; i < M1_FEAT; ++i) {
        dws[threadIdx.x] = dw1rx[i];
        __syncthreads();
        if (threadIdx.x < M1_H)
            out[M1_H * M1_H + 2 * M1_H + 1 + threadIdx.x * M1_FEAT + i] = dws[threadIdx.x] + dws[threadIdx.x + M1_H];
        __syncthreads();
    }
}

// ------------------------------------------------------- PE layer backward
// pe partial record per workgroup: dW1pe [128 * 64] (column 63 unused) | db1 [128]
#define M1_PEPART (M1_H * 64 + M1_H)

__global__ void __launch_bounds__(256) k_mlp1_pe_bwd(Mlp1Args a, const float* dpe1, float* dmu, float* part)
{
    __shared__ float wpe[M1_H * M1_KS];    // W1pe: wpe[u * M1_KS + c] = W1[u][c] (c < 63), 0 for c = 63
    __shared__ float dps[128 * M1_S];      // dpe1 rows of the tile
    __shared__ float pes[128 * M1_KS];     // PE rows of the tile; then dPE rows
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < M1_H * 64; i += 256) {
        const int u = i >> 6, c = i & 63;
        wpe[u * M1_KS + c] = c < M1_PE ? a.w1[u * M1_IN + c] : 0.0f;
    }
    floatx16 dw[2] = {floatx16{}, floatx16{}};    // dW1pe rows u = 32w.., columns c = 0..63
    float db1 = 0.0f;
    for (int tile = blockIdx.x; tile * 128 < a.P; tile += gridDim.x) {
        const int g0 = tile * 128;
        __syncthreads();
        for (int i = threadIdx.x; i < 128 * M1_H; i += 256) {
            const int m = i >> 7, u = i & (M1_H - 1), g = g0 + m;
            dps[m * M1_S + u] = g < a.P ? dpe1[(size_t)g * M1_H + u] : 0.0f;
        }
        pe_rows(a, g0, pes);
        __syncthreads();
        if (threadIdx.x < M1_H)
            for (int m = 0; m < 128; ++m) db1 += dps[m * M1_S + threadIdx.x];
        // dW1pe[u][c] += sum_m dpe1[m][u] PE[m][c]: wave w the rows u = 32w..; K = the 128 Gaussians
        {
            const float* arow = dps + (lane >> 5) * M1_S + 32 * w + (lane & 31);
            const float* bcol = pes + (lane >> 5) * M1_KS + (lane & 31);
            for (int ms = 0; ms < 64; ++ms) {
                const float av = arow[2 * ms * M1_S];
                dw[0] = mfma32x32x2(av, bcol[2 * ms * M1_KS], dw[0]);
                dw[1] = mfma32x32x2(av, bcol[2 * ms * M1_KS + 32], dw[1]);
            }
        }
        // dPE[m][c] = sum_u dpe1[m][u] W1pe[u][c]: wave w the Gaussians 32w..
        floatx16 dp[2] = {floatx16{}, floatx16{}};
        {
            const float* arow = dps + (32 * w + (lane & 31)) * M1_S + (lane >> 5);
            const float* bcol = wpe + (lane >> 5) * M1_KS + (lane & 31);
            for (int ks = 0; ks < M1_H / 2; ++ks) {
                const float av = arow[2 * ks];
                dp[0] = mfma32x32x2(av, bcol[2 * ks * M1_KS], dp[0]);
                dp[1] = mfma32x32x2(av, bcol[2 * ks * M1_KS + 32], dp[1]);
            }
        }
        __syncthreads();   // every wave's PE reads are done: dPE replaces PE in place
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int v = 0; v < 16; ++v) pes[(32 * w + acc_row(v, lane)) * M1_KS + 32 * t + (lane & 31)] = dp[t][v];
        __syncthreads();
        // the PE chain rule: thread (m, j) for Gaussian m, coordinate j
        for (int i = threadIdx.x; i < 128 * 3; i += 256) {
            const int m = i / 3, j = i % 3, g = g0 + m;
            if (g >= a.P) continue;
            const float* dpr = pes + m * M1_KS;
            const float mu = a.means[3 * (size_t)g + j];
            float s = dpr[j];
            for (int k = 0; k < 10; ++k) {
                float sn, cs;
                pe_sincos(mu, k, &sn, &cs);
                const float sc = 3.14159265358979323846f * (float)(1 << k);
                s += sc * (dpr[3 + 6 * k + j] * cs - dpr[6 + 6 * k + j] * sn);
            }
            dmu[3 * (size_t)g + j] += s;
        }
    }
    float* out = part + (size_t)blockIdx.x * M1_PEPART;
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int u = 32 * w + acc_row(v, lane), c = 32 * t + (lane & 31);
            out[u * 64 + c] = dw[t][v];
        }
    if (threadIdx.x < M1_H) out[M1_H * 64 + threadIdx.x] = db1;
}

// out[c] = sum over workgroups of part[wg][c] (record width `width`) in a
// fixed order: a workgroup takes 64 columns, its 4 waves the rows
// wg = slice mod 4 (each in increasing order, 4 rows in flight), and the
// four slice sums are added in slice order.
#define M1_CS_COLS 64
__global__ void __launch_bounds__(256) k_mlp1_colsum(const float* part, int nwg, int width, float* out)
{
    __shared__ float s_sum[4][M1_CS_COLS];
    const int l = threadIdx.x % M1_CS_COLS, sl = threadIdx.x / M1_CS_COLS;
    const int c = blockIdx.x * M1_CS_COLS + l;
    float s = 0.0f;
    if (c < width) {
        int i = sl;
        for (; i + 12 < nwg; i += 16) {
            const float v0 = part[(size_t)i * width + c], v1 = part[(size_t)(i + 4) * width + c];
            const float v2 = part[(size_t)(i + 8) * width + c], v3 = part[(size_t)(i + 12) * width + c];
            s += v0;
            s += v1;
            s += v2;
            s += v3;
        }
        for (; i < nwg; i += 4) s += part[(size_t)i * width + c];
    }
    s_sum[sl][l] = s;
    __syncthreads();
    if (sl == 0 && c < width) out[c] = ((s_sum[0][l] + s_sum[1][l]) + s_sum[2][l]) + s_sum[3][l];
}

// grads = dW1 [128 x 68] | db1 | dW2 [128 x 128] | db2 | dw3 | db3 from the
// two column sums
__global__ void __launch_bounds__(256) k_mlp1_assemble(const float* rowsum, const float* pesum, float* grads)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int nW1 = M1_H * M1_IN;
    if (i < nW1) {
        const int u = i / M1_IN, c = i % M1_IN;
        grads[i] = c < M1_PE ? pesum[u * 64 + c] : rowsum[M1_H * M1_H + 2 * M1_H + 1 + u * M1_FEAT + (c - M1_PE)];
    } else if (i < nW1 + M1_H) {
        grads[i] = pesum[M1_H * 64 + (i - nW1)];
    } else if (i < nW1 + M1_H + M1_H * M1_H) {
        grads[i] = rowsum[i - nW1 - M1_H];                                  // dW2
    } else if (i < nW1 + 2 * M1_H + M1_H * M1_H) {
        grads[i] = rowsum[M1_H * M1_H + M1_H + (i - nW1 - M1_H - M1_H * M1_H)];   // db2
    } else if (i < nW1 + 3 * M1_H + M1_H * M1_H) {
        grads[i] = rowsum[M1_H * M1_H + (i - nW1 - 2 * M1_H - M1_H * M1_H)];      // dw3
    } else if (i == nW1 + 3 * M1_H + M1_H * M1_H) {
        grads[i] = rowsum[M1_H * M1_H + 2 * M1_H];                           // db3
    }
}

__global__ void __launch_bounds__(256) k_mlp1_zero(float* p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = 0.0f;
}

// ---------------------------------------------------------------- launchers
static int grid_for(long long tiles) { return (int)(tiles < 1 ? 1 : (tiles < 1024 ? tiles : 1024)); }

int mlp1_bwd_groups(long long n_rows) { return grid_for((n_rows + 63) / 64); }
int mlp1_pe_groups(int P) { return grid_for(((long long)P + 127) / 128); }

size_t mlp1_work_floats(int P, int R)
{
    const long long rows = (long long)P * R;
    return 2 * (size_t)P * M1_H + (size_t)mlp1_bwd_groups(rows) * M1_PART + (size_t)mlp1_pe_groups(P) * M1_PEPART +
           M1_PART + M1_PEPART + 64;
}

size_t mlp1_grad_floats() { return (size_t)M1_H * M1_IN + 3 * M1_H + M1_H * M1_H + 1; }

hipError_t launch_mlp1_fwd(const Mlp1Args& a, hipStream_t s)
{
    if (a.P == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp1_pe, dim3(mlp1_pe_groups(a.P)), dim3(256), 0, s, a);
    const long long n_rows = (long long)a.P * a.R;
    if (n_rows == 0) return hipGetLastError();
    hipLaunchKernelGGL(k_mlp1_fwd, dim3(grid_for((n_rows + 127) / 128)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mlp1_bwd(const Mlp1Args& a, const float* dem, float* dmu, float* work, float* grads, hipStream_t s,
                           bool accumulate)
{
    const long long n_rows = (long long)a.P * a.R;
    float* dpe1 = work + (size_t)a.P * M1_H;
    const int nwg = mlp1_bwd_groups(n_rows), npe = mlp1_pe_groups(a.P);
    float* rpart = work + 2 * (size_t)a.P * M1_H;
    float* ppart = rpart + (size_t)nwg * M1_PART;
    float* rsum = ppart + (size_t)npe * M1_PEPART;
    float* psum = rsum + M1_PART;
    const size_t nz = (size_t)a.P * M1_H;
    hipLaunchKernelGGL(k_mlp1_zero, dim3((unsigned)std::max<size_t>(1, std::min<size_t>((nz + 255) / 256, 4096))),
                       dim3(256), 0, s, dpe1, nz);
    // the means gradient is added to (kernels below only add to dmu): cleared
    // first unless the caller accumulates into a gradient it already holds
    if (!accumulate)
        hipLaunchKernelGGL(k_mlp1_zero,
                           dim3((unsigned)std::max<size_t>(1, std::min<size_t>((3 * (size_t)a.P + 255) / 256, 4096))),
                           dim3(256), 0, s, dmu, 3 * (size_t)a.P);
    if (n_rows > 0) {
        hipLaunchKernelGGL(k_mlp1_bwd, dim3(nwg), dim3(256), 0, s, a, dem, dpe1, dmu, rpart);
        hipLaunchKernelGGL(k_mlp1_colsum, dim3((M1_PART + M1_CS_COLS - 1) / M1_CS_COLS), dim3(256), 0, s,
                           (const float*)rpart, nwg, (int)M1_PART, rsum);
    } else {
        hipLaunchKernelGGL(k_mlp1_zero, dim3(64), dim3(256), 0, s, rsum, (size_t)M1_PART);
    }
    if (a.P > 0) {
        hipLaunchKernelGGL(k_mlp1_pe_bwd, dim3(npe), dim3(256), 0, s, a, (const float*)dpe1, dmu, ppart);
        hipLaunchKernelGGL(k_mlp1_colsum, dim3((M1_PEPART + M1_CS_COLS - 1) / M1_CS_COLS), dim3(256), 0, s,
                           (const float*)ppart, npe, (int)M1_PEPART, psum);
    } else {
        hipLaunchKernelGGL(k_mlp1_zero, dim3(64), dim3(256), 0, s, psum, (size_t)M1_PEPART);
    }
    const size_t ng = mlp1_grad_floats();
    hipLaunchKernelGGL(k_mlp1_assemble, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, s, (const float*)rsum,
                       (const float*)psum, grads);
    return hipGetLastError();
}

}  // namespace gs
