// Shared emission MLP of the single-receiver training configuration on the
// matrix cores (FP32 MFMA, FP32 accumulate).  No counterpart in the reference
// (whose emission MLP is per Gaussian, 3->32->2, gaussian.py:47).
//
// Rows are (receiver, Gaussian) pairs, row = r * P + g.  With
//   h1 = relu(pe1[g] + W1rx rxf[row])      pe1 = PE(mu) W1pe^T + b1   [P,128]
//   h2 = relu(W2 h1 + b2)                  W2 [128,128] (out, in)
//   a  = w3 . h2 + b3
// the forward kernel computes a for 128-row tiles (4 waves x 32 rows; the
// grid strides over the tiles, so W2 is staged once per workgroup): h1 on
// the vector ALU into LDS, the 128x128 layer as 32x32x2 MFMA
// tiles (A = h1 rows from LDS, B = W2^T from LDS), relu, bias and the w3 dot
// in the epilogue with a 32-lane sum.  The backward kernel recomputes h1 and
// h2 for 64-row tiles and forms
//   g2  = da w3 (h2 > 0)                      (per row, in LDS)
//   dW2 += g2^T h1   (MFMA, K = rows; each wave keeps 32 rows of dW2)
//   dh1 = (g2 W2) (h1 > 0)                    (MFMA, written per row)
//   dw3 += da h2, db2 += g2, db3 += da
// with per-workgroup partial sums reduced by a column-sum kernel (fixed
// order).  What remains (the PE GEMM and its transpose, dW1rx = dh1^T rxf,
// the receiver-feature chain rule) is plain GEMM / elementwise work left to
// the caller.
//
// MFMA 32x32x2 f32 operand layout (wave64): lane l supplies A[l%32][l/32] and
// B[l/32][l%32]; accumulator element v of lane l is C[8(v/4) + 4(l/32) +
// v%4][l%32].  LDS rows are padded to 129 floats so that the 32 lanes of a
// half-wave hit 32 different banks in every operand read.
#include "gs_device.h"
#include "gs_kernels.h"

namespace gs {

#define M1_H 128     // hidden width
#define M1_S 129     // LDS row stride
#define M1_FEAT 5    // receiver features per row

typedef float floatx16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ floatx16 mfma32x32x2(float a, float b, floatx16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int acc_row(int v, int lane) { return 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3); }

// W2 [n][k] (global, row-major) -> w2t[k * M1_S + n]
__device__ __forceinline__ void load_w2t(const float* w2, float* w2t)
{
    for (int i = threadIdx.x; i < M1_H * M1_H; i += blockDim.x) {
        const int n = i >> 7, k = i & (M1_H - 1);
        w2t[k * M1_S + n] = w2[i];
    }
}

// h1 for `rows` rows starting at row0 into h1s[m * M1_S + u]; `threads`
// threads share the work (thread t: row t % rows, a slice of the units).
__device__ __forceinline__ void layer1(const Mlp1Args& a, long long row0, int rows, float* h1s)
{
    const long long n_rows = (long long)a.P * a.R;
    const int tid = threadIdx.x;
    const int m = tid % rows, part = tid / rows, nparts = blockDim.x / rows;
    const int per = M1_H / nparts, u0 = part * per;
    const long long row = row0 + m;
    float* dst = h1s + m * M1_S;
    if (row >= n_rows) {
        for (int u = u0; u < u0 + per; ++u) dst[u] = 0.0f;
        return;
    }
    const int g = (int)(row % a.P);
    float f[M1_FEAT];
#pragma unroll
    for (int i = 0; i < M1_FEAT; ++i) f[i] = a.rxf[row * M1_FEAT + i];
    const float* pe = a.pe1 + (size_t)g * M1_H;
    for (int u = u0; u < u0 + per; ++u) {
        const float* wr = a.w1rx + u * M1_FEAT;
        float v = pe[u];
#pragma unroll
        for (int i = 0; i < M1_FEAT; ++i) v = fmaf(wr[i], f[i], v);
        dst[u] = fmaxf(v, 0.0f);
    }
}

__global__ void __launch_bounds__(256) k_mlp1_fwd(Mlp1Args a)
{
    __shared__ float w2t[M1_H * M1_S];
    __shared__ float h1s[128 * M1_S];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long long n_rows = (long long)a.P * a.R;
    load_w2t(a.w2, w2t);   // once per workgroup: the grid strides over 128-row tiles
    const float b3 = a.b3[0];
    const float* hrow = h1s + (32 * w + (lane & 31)) * M1_S + (lane >> 5);
    const float* bcol = w2t + (lane >> 5) * M1_S + (lane & 31);
    for (long long tile = blockIdx.x; tile * 128 < n_rows; tile += gridDim.x) {
        const long long row0 = tile * 128;
        __syncthreads();   // w2t loaded / the previous tile's h1s reads done
        layer1(a, row0, 128, h1s);
        __syncthreads();
        floatx16 acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t] = floatx16{};
        for (int ks = 0; ks < M1_H / 2; ++ks) {
            const float av = hrow[2 * ks];
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = mfma32x32x2(av, bcol[2 * ks * M1_S + 32 * t], acc[t]);
        }
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
#pragma unroll
        for (int v = 0; v < 16; ++v)
#pragma unroll
            for (int off = 1; off < 32; off <<= 1) part[v] += __shfl_xor(part[v], off);
        if ((lane & 31) == 0) {
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const long long row = row0 + 32 * w + acc_row(v, lane);
                if (row < n_rows) {
                    a.out[2 * row] = part[v] + b3;
                    a.out[2 * row + 1] = 0.0f;
                }
            }
        }
    }
}

// partial record per workgroup: dW2 [128*128] | dw3 [128] | db2 [128] | db3
#define M1_PART (M1_H * M1_H + 2 * M1_H + 1)

__global__ void __launch_bounds__(256) k_mlp1_bwd(Mlp1Args a, const float* da, float* dh1, float* part)
{
    __shared__ float w2t[M1_H * M1_S];
    __shared__ float h1s[64 * M1_S];
    __shared__ float g2s[64 * M1_S];
    __shared__ float red[4][2][2][32];
    __shared__ float red3[256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int mh = w & 1, nh = w >> 1;     // this wave's row half and pair of 32-column tiles
    const long long n_rows = (long long)a.P * a.R;
    load_w2t(a.w2, w2t);
    floatx16 dw2[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) dw2[u] = floatx16{};
    float dw3p[2] = {0.0f, 0.0f}, db2p[2] = {0.0f, 0.0f}, db3p = 0.0f;
    for (long long tile = blockIdx.x; tile * 64 < n_rows; tile += gridDim.x) {
        const long long row0 = tile * 64;
        __syncthreads();   // the previous tile's LDS reads are done (and w2t is loaded)
        layer1(a, row0, 64, h1s);
        if (threadIdx.x < 64 && row0 + threadIdx.x < n_rows) db3p += da[2 * (row0 + threadIdx.x)];
        __syncthreads();
        // h2 for rows 32*mh.., columns of tiles 2*nh, 2*nh + 1
        floatx16 c[2] = {floatx16{}, floatx16{}};
        const float* hrow = h1s + (32 * mh + (lane & 31)) * M1_S + (lane >> 5);
        const float* bcol = w2t + (lane >> 5) * M1_S + 64 * nh + (lane & 31);
        for (int ks = 0; ks < M1_H / 2; ++ks) {
            const float av = hrow[2 * ks];
            c[0] = mfma32x32x2(av, bcol[2 * ks * M1_S], c[0]);
            c[1] = mfma32x32x2(av, bcol[2 * ks * M1_S + 32], c[1]);
        }
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int nn = 64 * nh + 32 * t + (lane & 31);
            const float bb = a.b2[nn], ww = a.w3[nn];
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int m = 32 * mh + acc_row(v, lane);
                const long long row = row0 + m;
                const float dav = row < n_rows ? da[2 * row] : 0.0f;
                const float h2 = fmaxf(c[t][v] + bb, 0.0f);
                const float g = h2 > 0.0f ? dav * ww : 0.0f;
                g2s[m * M1_S + nn] = g;
                dw3p[t] = fmaf(dav, h2, dw3p[t]);
                db2p[t] += g;
            }
        }
        __syncthreads();
        // dW2[n][k] += sum_m g2[m][n] h1[m][k]: this wave's n in 32*w..+32
        {
            const float* ga = g2s + (lane >> 5) * M1_S + 32 * w + (lane & 31);
            const float* hb = h1s + (lane >> 5) * M1_S + (lane & 31);
            for (int ms = 0; ms < 32; ++ms) {
                const float av = ga[2 * ms * M1_S];
#pragma unroll
                for (int u = 0; u < 4; ++u) dw2[u] = mfma32x32x2(av, hb[2 * ms * M1_S + 32 * u], dw2[u]);
            }
        }
        // dh1[m][k] = (sum_n g2[m][n] W2[n][k]) (h1 > 0): rows 32*mh.., column tiles 2*nh, 2*nh + 1
        {
            floatx16 e[2] = {floatx16{}, floatx16{}};
            const float* grow = g2s + (32 * mh + (lane & 31)) * M1_S + (lane >> 5);
            const float* wk = w2t + (64 * nh + (lane & 31)) * M1_S + (lane >> 5);
            for (int ns = 0; ns < M1_H / 2; ++ns) {
                const float av = grow[2 * ns];
                e[0] = mfma32x32x2(av, wk[2 * ns], e[0]);
                e[1] = mfma32x32x2(av, wk[32 * M1_S + 2 * ns], e[1]);
            }
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int k = 64 * nh + 32 * t + (lane & 31);
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int m = 32 * mh + acc_row(v, lane);
                    const long long row = row0 + m;
                    if (row < n_rows) dh1[row * M1_H + k] = h1s[m * M1_S + k] > 0.0f ? e[t][v] : 0.0f;
                }
            }
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
}

// out[c] = sum over workgroups of part[wg][c], in workgroup order
__global__ void __launch_bounds__(256) k_mlp1_colsum(const float* part, int nwg, float* out)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= M1_PART) return;
    float s = 0.0f;
    for (int i = 0; i < nwg; ++i) s += part[(size_t)i * M1_PART + c];
    out[c] = s;
}

int mlp1_bwd_groups(long long n_rows)
{
    const long long tiles = (n_rows + 63) / 64;
    return (int)(tiles < 1024 ? tiles : 1024);
}

size_t mlp1_part_floats(long long n_rows) { return (size_t)mlp1_bwd_groups(n_rows) * M1_PART; }

hipError_t launch_mlp1_fwd(const Mlp1Args& a, hipStream_t s)
{
    const long long n_rows = (long long)a.P * a.R;
    if (n_rows == 0) return hipSuccess;
    const long long tiles = (n_rows + 127) / 128;
    hipLaunchKernelGGL(k_mlp1_fwd, dim3((unsigned)(tiles < 1024 ? tiles : 1024)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mlp1_bwd(const Mlp1Args& a, const float* da, float* dh1, float* part, float* grads, hipStream_t s)
{
    const long long n_rows = (long long)a.P * a.R;
    if (n_rows == 0) return hipSuccess;
    const int nwg = mlp1_bwd_groups(n_rows);
    hipLaunchKernelGGL(k_mlp1_bwd, dim3(nwg), dim3(256), 0, s, a, da, dh1, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mlp1_colsum, dim3((M1_PART + 255) / 256), dim3(256), 0, s, part, nwg, grads);
    return hipGetLastError();
}

}  // namespace gs
