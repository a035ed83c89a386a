// Training loss of the multi-frequency (CSI) map, value and gradient.  No
// counterpart in the reference; it carries the reference's loss composition
// (train.py:418-427: (1 - lambda) * L1 + lambda * (1 - SSIM), lambda = 0.2 in
// optimizer.py:8) over to a complex map of GS_CSI_NC subcarriers:
//   L1    mean |out - target| over the real and imaginary parts of every
//         subcarrier of every pixel (R*N*NC*2 values);
//   SSIM  mean SSIM of the magnitude maps |C_k|, one 11x11-window SSIM plane
//         per (receiver, subcarrier) (R*NC planes of H x W).
// k_csi_mag turns the interleaved [R*N][NC][2] maps into magnitude planes
// [R*NC][N] and sums |out - target| per workgroup; the SSIM kernels run on the
// planes; k_csi_grad adds the L1 sign term and carries dL/d|C_k| back to
// (re, im) through C_k / |C_k| (0 at 0).  Both kernels take CL_PIX
// consecutive pixels per workgroup: the interleaved maps are read (and the
// gradient written) as consecutive 16-byte words of the block's contiguous
// span, and the magnitude planes go through an LDS transpose, so every global
// access is coalesced.
#include "gs_device.h"
#include "gs_kernels.h"

namespace gs {

#define CL_THREADS 256
#define CL_PIX 64                         // pixels per workgroup
#define CL_Q (GS_CSI_NC / 2)              // float4 per pixel (two subcarriers)

__global__ void __launch_bounds__(CL_THREADS) k_csi_mag(size_t npix, int N, const float* out, const float* tgt,
                                                        float* magA, float* magB, float* part)
{
    __shared__ float sA[GS_CSI_NC][CL_PIX + 1];
    __shared__ float sB[GS_CSI_NC][CL_PIX + 1];
    __shared__ float red[CL_THREADS];
    const size_t p0 = (size_t)blockIdx.x * CL_PIX;
    const int np = (int)min((size_t)CL_PIX, npix - p0);
    const float4* a = reinterpret_cast<const float4*>(out + p0 * (2 * GS_CSI_NC));
    const float4* b = reinterpret_cast<const float4*>(tgt + p0 * (2 * GS_CSI_NC));
    float l1 = 0.0f;
    // CL_PIX * CL_Q = 1664 words over 256 threads: up to 7 per thread, all
    // loaded before the first use (clamped to the block's span)
    constexpr int NI = (CL_PIX * CL_Q + CL_THREADS - 1) / CL_THREADS;
    float4 xs[NI], ys[NI];
#pragma unroll
    for (int u = 0; u < NI; ++u) {
        const int i = min((int)threadIdx.x + u * CL_THREADS, np * CL_Q - 1);
        xs[u] = a[i];
        ys[u] = b[i];
    }
#pragma unroll
    for (int u = 0; u < NI; ++u) {
        const int i = (int)threadIdx.x + u * CL_THREADS;
        if (i >= np * CL_Q) break;
        const int pl = i / CL_Q, k2 = i % CL_Q;
        const float4 x = xs[u], y = ys[u];
        sA[2 * k2][pl] = sqrtf(x.x * x.x + x.y * x.y);
        sA[2 * k2 + 1][pl] = sqrtf(x.z * x.z + x.w * x.w);
        sB[2 * k2][pl] = sqrtf(y.x * y.x + y.y * y.y);
        sB[2 * k2 + 1][pl] = sqrtf(y.z * y.z + y.w * y.w);
        l1 += fabsf(x.x - y.x) + fabsf(x.y - y.y) + fabsf(x.z - y.z) + fabsf(x.w - y.w);
    }
    red[threadIdx.x] = l1;
    __syncthreads();
    for (int i = threadIdx.x; i < GS_CSI_NC * CL_PIX; i += CL_THREADS) {
        const int k = i / CL_PIX, pl = i % CL_PIX;
        if (pl >= np) continue;
        const size_t p = p0 + pl, r = p / (size_t)N, q = p % (size_t)N;
        const size_t o = (r * GS_CSI_NC + k) * (size_t)N + q;
        magA[o] = sA[k][pl];
        magB[o] = sB[k][pl];
    }
    // fixed-order tree over the workgroup
    for (int w = CL_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__device__ __forceinline__ float sgnf(float v) { return v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f); }

__global__ void __launch_bounds__(CL_THREADS) k_csi_grad(size_t npix, int N, const float* out, const float* tgt,
                                                         const float* magA, const float* dmag, float l1w,
                                                         float* dout)
{
    __shared__ float sM[GS_CSI_NC][CL_PIX + 1];
    __shared__ float sD[GS_CSI_NC][CL_PIX + 1];
    const size_t p0 = (size_t)blockIdx.x * CL_PIX;
    const int np = (int)min((size_t)CL_PIX, npix - p0);
    {
        constexpr int NP = GS_CSI_NC * CL_PIX / CL_THREADS;
        static_assert(NP * CL_THREADS == GS_CSI_NC * CL_PIX, "whole planes per thread");
        float mv[NP], dv[NP];
#pragma unroll
        for (int u = 0; u < NP; ++u) {
            const int i = (int)threadIdx.x + u * CL_THREADS;
            const int k = i / CL_PIX, pl = min(i % CL_PIX, np - 1);
            const size_t p = p0 + pl, r = p / (size_t)N, q = p % (size_t)N;
            const size_t o = (r * GS_CSI_NC + k) * (size_t)N + q;
            mv[u] = magA[o];
            dv[u] = dmag[o];
        }
#pragma unroll
        for (int u = 0; u < NP; ++u) {
            const int i = (int)threadIdx.x + u * CL_THREADS;
            sM[i / CL_PIX][i % CL_PIX] = mv[u];
            sD[i / CL_PIX][i % CL_PIX] = dv[u];
        }
    }
    __syncthreads();
    const float4* a = reinterpret_cast<const float4*>(out + p0 * (2 * GS_CSI_NC));
    const float4* b = reinterpret_cast<const float4*>(tgt + p0 * (2 * GS_CSI_NC));
    float4* g = reinterpret_cast<float4*>(dout + p0 * (2 * GS_CSI_NC));
    constexpr int NI = (CL_PIX * CL_Q + CL_THREADS - 1) / CL_THREADS;
    float4 xs[NI], ys[NI];
#pragma unroll
    for (int u = 0; u < NI; ++u) {
        const int i = min((int)threadIdx.x + u * CL_THREADS, np * CL_Q - 1);
        xs[u] = a[i];
        ys[u] = b[i];
    }
#pragma unroll
    for (int u = 0; u < NI; ++u) {
        const int i = (int)threadIdx.x + u * CL_THREADS;
        if (i >= np * CL_Q) break;
        const int pl = i / CL_Q, k2 = i % CL_Q;
        const float4 x = xs[u], y = ys[u];
        const float m0 = sM[2 * k2][pl], m1 = sM[2 * k2 + 1][pl];
        const float f0 = m0 > 0.0f ? sD[2 * k2][pl] / m0 : 0.0f;
        const float f1 = m1 > 0.0f ? sD[2 * k2 + 1][pl] / m1 : 0.0f;
        g[i] = make_float4(fmaf(l1w, sgnf(x.x - y.x), f0 * x.x), fmaf(l1w, sgnf(x.y - y.y), f0 * x.y),
                           fmaf(l1w, sgnf(x.z - y.z), f1 * x.z), fmaf(l1w, sgnf(x.w - y.w), f1 * x.w));
    }
}

// Two-stage fixed-order reduction of the loss partials.  Stage 1: workgroup
// b of RD_BLOCKS sums a contiguous slice of the SSIM-map partials (ps, .x of
// each pair) and of the L1 partials (pl) in FP64; stage 2: one workgroup sums
// the RD_BLOCKS pairs and forms
//   loss = l1w * sum(l1) * inv_l + sw * (1 - sum(map) * inv_s).
#define RD_BLOCKS 256

__device__ __forceinline__ void block_sum2(double& a, double& b)
{
    __shared__ double r0[256], r1[256];
    const int tid = threadIdx.x;
    r0[tid] = a;
    r1[tid] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) {
            r0[tid] += r0[tid + w];
            r1[tid] += r1[tid + w];
        }
        __syncthreads();
    }
    a = r0[0];
    b = r1[0];
}

__global__ void __launch_bounds__(256) k_loss_stage1(const float* ps, int nbs, const float* pl, int nbl,
                                                     double* stage)
{
    const int b = blockIdx.x, tid = threadIdx.x;
    const int s0 = (int)((long long)nbs * b / RD_BLOCKS), s1 = (int)((long long)nbs * (b + 1) / RD_BLOCKS);
    const int l0 = (int)((long long)nbl * b / RD_BLOCKS), l1 = (int)((long long)nbl * (b + 1) / RD_BLOCKS);
    double m = 0.0, a = 0.0;
    for (int i = s0 + tid; i < s1; i += 256) m += (double)ps[2 * i];
    for (int i = l0 + tid; i < l1; i += 256) a += (double)pl[i];
    block_sum2(m, a);
    if (tid == 0) {
        stage[2 * b] = m;
        stage[2 * b + 1] = a;
    }
}

__global__ void __launch_bounds__(256) k_loss_stage2(const double* stage, double inv_s, double inv_l, float l1w,
                                                     float sw, float* loss)
{
    double m = stage[2 * threadIdx.x], a = stage[2 * threadIdx.x + 1];
    block_sum2(m, a);
    if (threadIdx.x == 0) loss[0] = (float)((double)l1w * a * inv_l + (double)sw * (1.0 - m * inv_s));
}

int csi_loss_blocks(size_t npix) { return (int)((npix + CL_PIX - 1) / CL_PIX); }

hipError_t launch_csi_mag(size_t npix, int N, const float* out, const float* tgt, float* magA, float* magB,
                          float* part, hipStream_t s)
{
    if (npix == 0) return hipSuccess;
    hipLaunchKernelGGL(k_csi_mag, dim3(csi_loss_blocks(npix)), dim3(CL_THREADS), 0, s, npix, N, out, tgt, magA, magB,
                       part);
    return hipGetLastError();
}

hipError_t launch_csi_grad(size_t npix, int N, const float* out, const float* tgt, const float* magA,
                           const float* dmag, float l1w, float* dout, hipStream_t s)
{
    if (npix == 0) return hipSuccess;
    hipLaunchKernelGGL(k_csi_grad, dim3(csi_loss_blocks(npix)), dim3(CL_THREADS), 0, s, npix, N, out, tgt, magA,
                       dmag, l1w, dout);
    return hipGetLastError();
}

size_t loss_reduce2_scratch() { return 2 * RD_BLOCKS * sizeof(double); }

hipError_t launch_loss_reduce2(const float* ps, int nbs, double inv_s, const float* pl, int nbl, double inv_l,
                               float l1w, float sw, double* stage, float* loss, hipStream_t s)
{
    hipLaunchKernelGGL(k_loss_stage1, dim3(RD_BLOCKS), dim3(256), 0, s, ps, nbs, pl, nbl, stage);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_loss_stage2, dim3(1), dim3(RD_BLOCKS), 0, s, stage, inv_s, inv_l, l1w, sw, loss);
    return hipGetLastError();
}

}  // namespace gs
