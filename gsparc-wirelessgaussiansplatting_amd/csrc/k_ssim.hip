// Fused SSIM forward and backward (fused-ssim/ssim.cu:187-366) over a batch
// of planes (B * CH planes of H x W; for the rasterizer, one plane per
// receiver spectrum).
//
// A 256-thread workgroup owns a 32 x 32 output tile.  The zero-padded halo
// (42 x 42) of the inputs is staged in LDS once; the horizontal 11-tap pass
// produces all five (forward) or three (backward) filtered quantities for
// every halo row, then the vertical pass finishes them per output pixel.
// Both passes slide along registers: a thread filters a run of SS_RUN
// consecutive outputs (along a row, then down a column) from the SS_RUN + 10
// inputs it reads once from LDS, and the forward's squares and cross products
// are formed once per input rather than once per tap.  Summation order
// matches the reference (taps 0..10, horizontal first; every output's own
// chain of fused multiply-adds, as CUDA compiles the reference's
// `acc += w * x`), and the per-pixel partial derivatives share the two
// reciprocals 1/A and 1/B, so the values agree with the CPU restatement to
// rounding (a few ulp).
#include <cmath>

#include "gs_blend.h"
#include "gs_device.h"
#include "gs_kernels.h"

namespace gs {

#define SS_TW 32
#define SS_TH 32
#define SS_HW (SS_TW + 10)
#define SS_HH (SS_TH + 10)
#define SS_THREADS 256
#define SS_RUN 4                                // outputs per register run
#define SS_HRUNS (SS_HH * (SS_TW / SS_RUN))     // horizontal runs per tile
static_assert((SS_TH / SS_RUN) * SS_TW == SS_THREADS, "one vertical run per thread");

struct Taps { float w[11]; };

// The tile of this workgroup.  Workgroups are dealt to the 8 XCDs round
// robin by id; tile t goes to the workgroup that gives XCD x a contiguous
// range of tiles, so that neighbouring tiles, which share halo pixels, share
// that XCD's L2.
struct SsTile { int x, y, z; };

__device__ __forceinline__ SsTile ssim_tile()
{
    const int gx = gridDim.x, gy = gridDim.y;
    const int n = gx * gy * gridDim.z;
    const int L = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const int xcd = L & 7, k = L >> 3, per = n >> 3, rem = n & 7;
    const int t = xcd < rem ? xcd * (per + 1) + k : rem * (per + 1) + (xcd - rem) * per + k;
    SsTile r;
    r.x = t % gx;
    r.y = (t / gx) % gy;
    r.z = t / (gx * gy);
    return r;
}

__global__ void __launch_bounds__(SS_THREADS) k_ssim_fwd(SsimArgs a, Taps tp)
{
    __shared__ float s1[SS_HH][SS_HW + 1];
    __shared__ float s2[SS_HH][SS_HW + 1];
    __shared__ float h[5][SS_HH][SS_TW + 1];
    const int tid = threadIdx.x;
    const SsTile tl = ssim_tile();
    const int x0 = tl.x * SS_TW, y0 = tl.y * SS_TH;
    const size_t plane = (size_t)tl.z * a.H * a.W;
    const float* im1 = a.img1 + plane;
    const float* im2 = a.img2 + plane;
    // halo staging: all of a thread's loads are issued before the first LDS
    // store (clamped in-plane addresses, zero outside), so that their
    // latencies overlap
    {
        constexpr int NS = (SS_HH * SS_HW + SS_THREADS - 1) / SS_THREADS;
        float v1[NS], v2[NS];
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int i = tid + u * SS_THREADS;
            const int ly = i / SS_HW, lx = i % SS_HW;
            const int gy = y0 + ly - 5, gx = x0 + lx - 5;
            const size_t o = (size_t)min(max(gy, 0), a.H - 1) * a.W + min(max(gx, 0), a.W - 1);
            v1[u] = im1[o];
            v2[u] = im2[o];
        }
        __builtin_amdgcn_sched_barrier(0);   // keep the loads ahead of the stores
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int i = tid + u * SS_THREADS;
            if (i >= SS_HH * SS_HW) break;
            const int ly = i / SS_HW, lx = i % SS_HW;
            const int gy = y0 + ly - 5, gx = x0 + lx - 5;
            const bool in = gx >= 0 && gx < a.W && gy >= 0 && gy < a.H;
            s1[ly][lx] = in ? v1[u] : 0.0f;
            s2[ly][lx] = in ? v2[u] : 0.0f;
        }
    }
    __syncthreads();
    // horizontal: run t covers halo row ly, output columns c0 .. c0 + SS_RUN - 1
    for (int t = tid; t < SS_HRUNS; t += SS_THREADS) {
        const int ly = t / (SS_TW / SS_RUN), c0 = (t % (SS_TW / SS_RUN)) * SS_RUN;
        float p[SS_RUN + 10], q[SS_RUN + 10];
#pragma unroll
        for (int i = 0; i < SS_RUN + 10; ++i) {
            p[i] = s1[ly][c0 + i];
            q[i] = s2[ly][c0 + i];
        }
        float pp[SS_RUN + 10], qq[SS_RUN + 10], pq[SS_RUN + 10];
#pragma unroll
        for (int i = 0; i < SS_RUN + 10; ++i) {
            pp[i] = p[i] * p[i];
            qq[i] = q[i] * q[i];
            pq[i] = p[i] * q[i];
        }
#pragma unroll
        for (int o = 0; o < SS_RUN; ++o) {
            float m1 = 0.0f, m2 = 0.0f, e11 = 0.0f, e22 = 0.0f, e12 = 0.0f;
#pragma unroll
            for (int k = 0; k < 11; ++k) {
                m1 = fmaf(tp.w[k], p[o + k], m1);
                m2 = fmaf(tp.w[k], q[o + k], m2);
                e11 = fmaf(tp.w[k], pp[o + k], e11);
                e22 = fmaf(tp.w[k], qq[o + k], e22);
                e12 = fmaf(tp.w[k], pq[o + k], e12);
            }
            h[0][ly][c0 + o] = m1; h[1][ly][c0 + o] = m2; h[2][ly][c0 + o] = e11;
            h[3][ly][c0 + o] = e22; h[4][ly][c0 + o] = e12;
        }
    }
    __syncthreads();
    // vertical: the thread's column lx, output rows r0 .. r0 + SS_RUN - 1
    const int lx = tid % SS_TW, r0 = (tid / SS_TW) * SS_RUN;
    float v[SS_RUN][5];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        float col[SS_RUN + 10];
#pragma unroll
        for (int i = 0; i < SS_RUN + 10; ++i) col[i] = h[c][r0 + i][lx];
#pragma unroll
        for (int o = 0; o < SS_RUN; ++o) {
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < 11; ++k) acc = fmaf(tp.w[k], col[o + k], acc);
            v[o][c] = acc;
        }
    }
    float sum_map = 0.0f, sum_abs = 0.0f, sum_d2 = 0.0f, sum_p2 = 0.0f, sum_t2 = 0.0f;
#pragma unroll
    for (int o = 0; o < SS_RUN; ++o) {
        const int ly = r0 + o;
        const int gy = y0 + ly, gx = x0 + lx;
        if (gy >= a.H || gx >= a.W) continue;
        const float mu1 = v[o][0], mu2 = v[o][1];
        const float sigma1_sq = v[o][2] - mu1 * mu1;
        const float sigma2_sq = v[o][3] - mu2 * mu2;
        const float sigma12 = v[o][4] - mu1 * mu2;
        const float C = 2.0f * mu1 * mu2 + a.C1;
        const float D = 2.0f * sigma12 + a.C2;
        const float A = mu1 * mu1 + mu2 * mu2 + a.C1;
        const float B = sigma1_sq + sigma2_sq + a.C2;
        const size_t o_ = plane + (size_t)gy * a.W + gx;
        // ssim.cu:269-283 with the divisions shared: iA = 1/A, iB = 1/B
        const float iA = 1.0f / A, iB = 1.0f / B, iAB = iA * iB;
        const float CD = C * D;
        const float m = CD * iAB;
        if (a.map) a.map[o_] = m;
        if (a.dmu1) {
            // (2 mu2 D - 2 mu2 C) / (AB) - 2 mu1 CD / (A^2 B) + 2 mu1 CD / (A B^2)
            a.dmu1[o_] = 2.0f * iAB * fmaf(mu1 * CD, iB - iA, mu2 * (D - C));
            a.ds1[o_] = -CD * iAB * iB;
            a.ds12[o_] = 2.0f * C * iAB;
        }
        sum_map += m;
        const float p = s1[ly + 5][lx + 5], t = s2[ly + 5][lx + 5];   // the staged pixel values
        sum_abs += fabsf(p - t);
        if (a.nsum == 5) {
            sum_d2 += (p - t) * (p - t);
            sum_p2 += p * p;
            sum_t2 += t * t;
        }
    }
    if (a.part) {
        // fixed-order sums over the workgroup (the loss is the same run to
        // run): each wave's sum by lane exchanges in a fixed pattern, then the
        // waves' sums in wave order
        constexpr int NWV = SS_THREADS / GS_WAVE;
        __shared__ float r[5][NWV];
        const int ns = a.nsum;
        const float vals[5] = {sum_map, sum_abs, sum_d2, sum_p2, sum_t2};
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            if (c >= ns) break;   // uniform
            const float w = wave_sum(vals[c]);
            if (tid % GS_WAVE == 0) r[c][tid / GS_WAVE] = w;
        }
        __syncthreads();
        if (tid == 0) {
            const size_t b = ((size_t)tl.z * gridDim.y + tl.y) * gridDim.x + tl.x;
            for (int c = 0; c < ns; ++c) {
                float t = r[c][0];
#pragma unroll
                for (int w = 1; w < NWV; ++w) t += r[c][w];
                a.part[(size_t)ns * b + c] = t;
            }
        }
    }
}

__global__ void __launch_bounds__(SS_THREADS) k_ssim_bwd(SsimBwdArgs a, Taps tp)
{
    __shared__ float s[3][SS_HH][SS_HW + 1];
    __shared__ float h[3][SS_HH][SS_TW + 1];
    const int tid = threadIdx.x;
    const SsTile tl = ssim_tile();
    const int x0 = tl.x * SS_TW, y0 = tl.y * SS_TH;
    const size_t plane = (size_t)tl.z * a.H * a.W;
    // the plane's loss weight (the confidence of its spectrum), wave-uniform
    const float ps = a.plane_scale ? a.plane_scale[tl.z] : 1.0f;
    const float dmap_const = a.dmap_const * ps, l1w = a.l1w * ps;
    // the images at the thread's output pixels (column lx, rows r0 ..
    // r0 + SS_RUN - 1 of the tile), loaded now and used at the end
    const int lx = tid % SS_TW, r0 = (tid / SS_TW) * SS_RUN;
    float xi[SS_RUN], yi[SS_RUN];
#pragma unroll
    for (int u = 0; u < SS_RUN; ++u) {
        const int gy = min(y0 + r0 + u, a.H - 1), gx = min(x0 + lx, a.W - 1);
        xi[u] = a.img1[plane + (size_t)gy * a.W + gx];
        yi[u] = a.img2[plane + (size_t)gy * a.W + gx];
    }
    {
        // all of a thread's loads first (clamped in-plane addresses, zero
        // outside), so that their latencies overlap
        constexpr int NS = (SS_HH * SS_HW + SS_THREADS - 1) / SS_THREADS;
        float g[NS], t0[NS], t1[NS], t2[NS];
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int i = min(tid + u * SS_THREADS, SS_HH * SS_HW - 1);
            const int ly = i / SS_HW, lx_ = i % SS_HW;
            const int gy = y0 + ly - 5, gx = x0 + lx_ - 5;
            const size_t o = plane + (size_t)min(max(gy, 0), a.H - 1) * a.W + min(max(gx, 0), a.W - 1);
            g[u] = a.dmap ? a.dmap[o] : dmap_const;
            t0[u] = a.dmu1[o];
            t1[u] = a.ds1[o];
            t2[u] = a.ds12[o];
        }
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int i = tid + u * SS_THREADS;
            if (i >= SS_HH * SS_HW) break;
            const int ly = i / SS_HW, lx_ = i % SS_HW;
            const int gy = y0 + ly - 5, gx = x0 + lx_ - 5;
            const bool in = gx >= 0 && gx < a.W && gy >= 0 && gy < a.H;
            s[0][ly][lx_] = in ? g[u] * t0[u] : 0.0f;
            s[1][ly][lx_] = in ? g[u] * t1[u] : 0.0f;
            s[2][ly][lx_] = in ? g[u] * t2[u] : 0.0f;
        }
    }
    __syncthreads();
    // horizontal: run t covers halo row ly, output columns c0 .. c0 + SS_RUN - 1
    for (int t = tid; t < SS_HRUNS; t += SS_THREADS) {
        const int ly = t / (SS_TW / SS_RUN), c0 = (t % (SS_TW / SS_RUN)) * SS_RUN;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float row[SS_RUN + 10];
#pragma unroll
            for (int i = 0; i < SS_RUN + 10; ++i) row[i] = s[c][ly][c0 + i];
#pragma unroll
            for (int o = 0; o < SS_RUN; ++o) {
                float acc = 0.0f;
#pragma unroll
                for (int k = 0; k < 11; ++k) acc = fmaf(tp.w[k], row[o + k], acc);
                h[c][ly][c0 + o] = acc;
            }
        }
    }
    __syncthreads();
    float v[SS_RUN][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float col[SS_RUN + 10];
#pragma unroll
        for (int i = 0; i < SS_RUN + 10; ++i) col[i] = h[c][r0 + i][lx];
#pragma unroll
        for (int o = 0; o < SS_RUN; ++o) {
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < 11; ++k) acc = fmaf(tp.w[k], col[o + k], acc);
            v[o][c] = acc;
        }
    }
#pragma unroll
    for (int u = 0; u < SS_RUN; ++u) {
        const int gy = y0 + r0 + u, gx = x0 + lx;
        if (gy >= a.H || gx >= a.W) continue;
        const size_t o = plane + (size_t)gy * a.W + gx;
        const float x = xi[u], y = yi[u];
        float g = v[u][0] + x * 2.0f * v[u][1] + y * v[u][2];
        if (l1w != 0.0f) {
            // d|x - y|/dx with sign(0) = 0, as torch's abs backward
            const float sgn = x > y ? 1.0f : (x < y ? -1.0f : 0.0f);
            g = g + l1w * sgn;
        }
        a.dimg1[o] = g;
    }
}

__global__ void __launch_bounds__(256) k_loss_reduce(const float* part, int nb, double inv_n, float l1w, float sw,
                                                     float* loss)
{
    __shared__ double r0[256], r1[256];
    const int tid = threadIdx.x;
    double m = 0.0, a = 0.0;
    for (int b = tid; b < nb; b += 256) {
        m += (double)part[2 * b];
        a += (double)part[2 * b + 1];
    }
    r0[tid] = m;
    r1[tid] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) {
            r0[tid] += r0[tid + w];
            r1[tid] += r1[tid + w];
        }
        __syncthreads();
    }
    if (tid == 0) loss[0] = (float)((double)l1w * r1[0] * inv_n + (double)sw * (1.0 - r0[0] * inv_n));
}

// One workgroup.  Phase 1, per plane (threads strided over the planes): the
// plane's L1 and SSIM sums from its workgroups' partials (in workgroup
// order, in FP64), its reconstruction loss L_r, its confidence C_r (the
// ConfidenceMLP of gaussian.py:575-609, FP32 dot products in index order)
// and its loss term; the terms are summed by a fixed tree.  Phase 2, per
// hidden unit (thread j < hidden, over the planes in order): the gradient of
// the mean loss with respect to the network, through dloss/dC_r =
// (L_r - alpha / C_r) / R and dC/draw = C - 1.
// C = 1 + exp(w2 . relu(W1 x + b1) + b2) of the packed confidence network
// [W1 (hidden x 3) | b1 | w2 | b2] at receiver position x (gaussian.py:598-609)
__device__ float conf_value(const float* mlp, int Hn, const float* x3)
{
    const float* W1 = mlp;
    const float* b1 = mlp + 3 * Hn;
    const float* w2 = mlp + 4 * Hn;
    const float x = x3[0], y = x3[1], z = x3[2];
    float raw = 0.0f;
    for (int j = 0; j < Hn; ++j) {
        float pre = fmaf(W1[3 * j], x, b1[j]);
        pre = fmaf(W1[3 * j + 1], y, pre);
        pre = fmaf(W1[3 * j + 2], z, pre);
        raw = fmaf(w2[j], fmaxf(pre, 0.0f), raw);
    }
    return 1.0f + expf(raw + mlp[5 * Hn]);
}

// The evaluation metrics of every plane (train.py:76-102, 651-693) from the
// five per-workgroup sums of k_ssim_fwd: loss, PSNR, SSIM, MAE, MSE,
// RSSI of the rendering, RSSI difference, confidence.  One thread per plane;
// fixed-order sums in double.
__global__ void __launch_bounds__(256) k_eval_metrics(EvalArgs a)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= a.planes) return;
    double sm = 0.0, sa = 0.0, sd2 = 0.0, sp2 = 0.0, st2 = 0.0;
    const float* pp = a.part + 5 * (size_t)r * a.nbp;
    for (int b = 0; b < a.nbp; ++b) {
        sm += (double)pp[5 * b];
        sa += (double)pp[5 * b + 1];
        sd2 += (double)pp[5 * b + 2];
        sp2 += (double)pp[5 * b + 3];
        st2 += (double)pp[5 * b + 4];
    }
    const double ssim = sm * a.inv_np, mae = sa * a.inv_np, mse = sd2 * a.inv_np;
    const double rssi_p = 10.0 * log10(sp2 * a.inv_np) - 30.0;
    const double rssi_t = 10.0 * log10(st2 * a.inv_np) - 30.0;
    float* out = a.metrics + (size_t)EVAL_METRICS * r;
    out[0] = (float)((1.0 - (double)a.lambda_ssim) * mae + (double)a.lambda_ssim * (1.0 - ssim));
    out[1] = (float)(20.0 * log10(1.0 / sqrt(mse)));
    out[2] = (float)ssim;
    out[3] = (float)mae;
    out[4] = (float)mse;
    out[5] = (float)rssi_p;
    out[6] = (float)(rssi_p - rssi_t);
    out[7] = a.mlp ? conf_value(a.mlp, a.hidden, a.rx + 3 * (size_t)r) : 1.0f;
}

hipError_t launch_eval_metrics(const EvalArgs& a, hipStream_t s)
{
    if (a.planes == 0) return hipSuccess;
    hipLaunchKernelGGL(k_eval_metrics, dim3((unsigned)((a.planes + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

#define CONF_CHUNK 256   // planes staged in LDS at a time by the weight gradients
__global__ void __launch_bounds__(256) k_conf_loss(ConfLossArgs a)
{
    __shared__ double red[256];
    __shared__ float s_draw[CONF_CHUNK];
    __shared__ float s_rx[CONF_CHUNK * 3];
    __shared__ double s_g[256 * 5];
    const int tid = threadIdx.x, Hn = a.hidden;
    const float* W1 = a.mlp;
    const float* b1 = a.mlp ? a.mlp + 3 * Hn : nullptr;
    const float* w2 = a.mlp ? a.mlp + 4 * Hn : nullptr;
    const float b2 = a.mlp ? a.mlp[5 * Hn] : 0.0f;
    const double invR = 1.0 / (double)a.planes;
    double term = 0.0;
    for (int r = tid; r < a.planes; r += 256) {
        double m = 0.0, l1 = 0.0;
        const float* pp = a.part + 2 * (size_t)r * a.nbp;
#pragma unroll 4
        for (int b = 0; b < a.nbp; ++b) {
            m += (double)pp[2 * b];
            l1 += (double)pp[2 * b + 1];
        }
        const double L = (double)a.l1w * l1 * a.inv_np + (double)a.sw * (1.0 - m * a.inv_np);
        float C = 1.0f;
        if (a.mlp) {
            const float x = a.rx[3 * r], y = a.rx[3 * r + 1], z = a.rx[3 * r + 2];
            float raw = 0.0f;
            for (int j = 0; j < Hn; ++j) {
                float pre = fmaf(W1[3 * j], x, b1[j]);
                pre = fmaf(W1[3 * j + 1], y, pre);
                pre = fmaf(W1[3 * j + 2], z, pre);
                raw = fmaf(w2[j], fmaxf(pre, 0.0f), raw);
            }
            C = 1.0f + expf(raw + b2);
        }
        term += (double)C * L - (double)a.alpha * log((double)C);
        a.plane_scale[r] = C;
        if (a.rec) a.rec[r] = (float)L;
        if (a.conf) a.conf[r] = C;
        // dloss / draw_r for phase 2
        a.scratch[4 * r] = (float)(((L - (double)a.alpha / (double)C) * invR) * ((double)C - 1.0));
    }
    red[tid] = term;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) red[tid] += red[tid + w];
        __syncthreads();
    }
    if (tid == 0) a.loss[0] = (float)(red[0] * invR);
    if (!a.mlp || !a.dmlp) return;
    __threadfence();
    __syncthreads();
    // weight gradients: hidden unit j = tid % Hn over the planes of slice
    // tid / Hn (S = 256 / Hn slices, planes i = slice, slice + S, ... of each
    // LDS chunk), the slices then combined in slice order
    const int S = 256 / Hn;
    const int j = tid % Hn, sl = tid / Hn;
    const bool act = sl < S;
    float wx = 0.0f, wy = 0.0f, wz = 0.0f, bj = 0.0f, vj = 0.0f;
    if (act) {
        wx = W1[3 * j]; wy = W1[3 * j + 1]; wz = W1[3 * j + 2];
        bj = b1[j]; vj = w2[j];
    }
    double gw0 = 0.0, gw1 = 0.0, gw2 = 0.0, gb = 0.0, g2 = 0.0, gb2 = 0.0;
    for (int c0 = 0; c0 < a.planes; c0 += CONF_CHUNK) {
        const int nc = min(CONF_CHUNK, a.planes - c0);
        __syncthreads();   // the previous chunk is consumed
        for (int i = tid; i < nc; i += 256) {
            s_draw[i] = a.scratch[4 * (c0 + i)];
            s_rx[3 * i] = a.rx[3 * (c0 + i)];
            s_rx[3 * i + 1] = a.rx[3 * (c0 + i) + 1];
            s_rx[3 * i + 2] = a.rx[3 * (c0 + i) + 2];
        }
        __syncthreads();
        if (act) {
            for (int i = sl; i < nc; i += S) {
                const float x = s_rx[3 * i], y = s_rx[3 * i + 1], z = s_rx[3 * i + 2];
                float pre = fmaf(wx, x, bj);
                pre = fmaf(wy, y, pre);
                pre = fmaf(wz, z, pre);
                const double draw = (double)s_draw[i];
                g2 += draw * (double)fmaxf(pre, 0.0f);
                const double dh = pre > 0.0f ? draw * (double)vj : 0.0;
                gw0 += dh * x;
                gw1 += dh * y;
                gw2 += dh * z;
                gb += dh;
            }
        }
        if (tid == 0)
            for (int i = 0; i < nc; ++i) gb2 += (double)s_draw[i];   // planes in order
    }
    if (act) {
        s_g[5 * tid] = gw0;
        s_g[5 * tid + 1] = gw1;
        s_g[5 * tid + 2] = gw2;
        s_g[5 * tid + 3] = gb;
        s_g[5 * tid + 4] = g2;
    }
    __syncthreads();
    if (tid < Hn) {
        double t[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int q = 0; q < S; ++q)
#pragma unroll
            for (int k = 0; k < 5; ++k) t[k] += s_g[5 * (q * Hn + tid) + k];
        a.dmlp[3 * tid] = (float)t[0];
        a.dmlp[3 * tid + 1] = (float)t[1];
        a.dmlp[3 * tid + 2] = (float)t[2];
        a.dmlp[3 * Hn + tid] = (float)t[3];
        a.dmlp[4 * Hn + tid] = (float)t[4];
    }
    if (tid == 0) a.dmlp[5 * Hn] = (float)gb2;
}

hipError_t launch_conf_loss(const ConfLossArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(k_conf_loss, dim3(1), dim3(256), 0, s, a);
    return hipGetLastError();
}

// 11-tap Gaussian window, sigma 1.5 (utils.py:116-124): float32 samples
// normalised by their sum, which reproduces the float literals of
// ssim.cu:9-19 bit for bit.
static Taps make_taps()
{
    Taps t;
    double s = 0.0;
    for (int i = 0; i < 11; ++i) {
        t.w[i] = (float)std::exp(-(double)((i - 5) * (i - 5)) / (2.0 * 1.5 * 1.5));
        s += t.w[i];
    }
    const float sf = (float)s;
    for (int i = 0; i < 11; ++i) t.w[i] = t.w[i] / sf;
    return t;
}

hipError_t launch_ssim_fwd(const SsimArgs& a, hipStream_t s)
{
    if (a.planes == 0 || a.H == 0 || a.W == 0) return hipSuccess;
    static const Taps tp = make_taps();
    dim3 grid((a.W + SS_TW - 1) / SS_TW, (a.H + SS_TH - 1) / SS_TH, a.planes);
    hipLaunchKernelGGL(k_ssim_fwd, grid, dim3(SS_THREADS), 0, s, a, tp);
    return hipGetLastError();
}

int ssim_blocks(int planes, int H, int W)
{
    return ((W + SS_TW - 1) / SS_TW) * ((H + SS_TH - 1) / SS_TH) * planes;
}

hipError_t launch_loss_reduce(const float* part, int nb, double inv_n, float l1w, float sw, float* loss,
                              hipStream_t s)
{
    hipLaunchKernelGGL(k_loss_reduce, dim3(1), dim3(256), 0, s, part, nb, inv_n, l1w, sw, loss);
    return hipGetLastError();
}

hipError_t launch_ssim_bwd(const SsimBwdArgs& a, hipStream_t s)
{
    if (a.planes == 0 || a.H == 0 || a.W == 0) return hipSuccess;
    static const Taps tp = make_taps();
    dim3 grid((a.W + SS_TW - 1) / SS_TW, (a.H + SS_TH - 1) / SS_TH, a.planes);
    hipLaunchKernelGGL(k_ssim_bwd, grid, dim3(SS_THREADS), 0, s, a, tp);
    return hipGetLastError();
}

}  // namespace gs
