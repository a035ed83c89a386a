// The spectrum dataset on the device (preprocess_data.py:114-203): the view
// matrix of every receiver and the gather of an iteration's batch.
#include <cmath>

#include "gs_device.h"
#include "gs_kernels.h"

namespace gs {

// build_view_matrix (preprocess_data.py:177-202): R_c2w from the gateway
// quaternion (x, y, z, w; normalised, as scipy's Rotation.from_quat and
// as_matrix compute it in double, then rounded to float), R_w2c = R_c2w^T,
// t = -(R_w2c C) in float, V = [R_w2c t; 0 1] stored transposed (column-major,
// the layout the rasterizer reads).  out: [N, 16].
__global__ void __launch_bounds__(256) k_view_matrices(int N, double qx, double qy, double qz, double qw,
                                                       const float* rx, float* out)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    const double n = sqrt(qx * qx + qy * qy + qz * qz + qw * qw);
    const double x = qx / n, y = qy / n, z = qz / n, w = qw / n;
    const double x2 = x * x, y2 = y * y, z2 = z * z, w2 = w * w;
    const double xy = x * y, zw = z * w, xz = x * z, yw = y * w, yz = y * z, xw = x * w;
    double Rd[3][3];
    Rd[0][0] = x2 - y2 - z2 + w2;
    Rd[1][0] = 2.0 * (xy + zw);
    Rd[2][0] = 2.0 * (xz - yw);
    Rd[0][1] = 2.0 * (xy - zw);
    Rd[1][1] = -x2 + y2 - z2 + w2;
    Rd[2][1] = 2.0 * (yz + xw);
    Rd[0][2] = 2.0 * (xz + yw);
    Rd[1][2] = 2.0 * (yz - xw);
    Rd[2][2] = -x2 - y2 + z2 + w2;
    float Rw2c[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) Rw2c[a][b] = (float)Rd[b][a];
    const float c0 = rx[3 * (size_t)i], c1 = rx[3 * (size_t)i + 1], c2 = rx[3 * (size_t)i + 2];
    float t[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) t[a] = -((Rw2c[a][0] * c0 + Rw2c[a][1] * c1) + Rw2c[a][2] * c2);
    float* o = out + 16 * (size_t)i;
    // o[4 j + k] = V[k][j]
#pragma unroll
    for (int j = 0; j < 3; ++j) {
#pragma unroll
        for (int k = 0; k < 3; ++k) o[4 * j + k] = Rw2c[k][j];
        o[4 * j + 3] = 0.0f;
    }
    o[12] = t[0];
    o[13] = t[1];
    o[14] = t[2];
    o[15] = 1.0f;
}

// The batch of the iteration in train_state[0] (1-based): sample r is
// order[((iteration - 1) R + r) mod L]; its spectrum [npix], receiver
// position [3] and view matrix [16] are copied into the batch buffers.
// One workgroup row per sample; float4 copies when npix allows.
__global__ void __launch_bounds__(256) k_gather_batch(int R, int N, int npix, const int* order, int L,
                                                      const uint32_t* train_state, const float* spectra,
                                                      const float* rx, const float* views, float* out_spec,
                                                      float* out_rx, float* out_views)
{
    const int r = blockIdx.y;
    const long long it = (long long)train_state[0];
    const long long pos = (((it - 1) * R + r) % L + L) % L;
    int src = order[pos];
    src = src < 0 ? 0 : (src >= N ? N - 1 : src);
    const float* s = spectra + (size_t)src * npix;
    float* d = out_spec + (size_t)r * npix;
    if ((npix & 3) == 0) {
        const float4* s4 = reinterpret_cast<const float4*>(s);
        float4* d4 = reinterpret_cast<float4*>(d);
        for (int k = blockIdx.x * 256 + threadIdx.x; k < npix / 4; k += gridDim.x * 256) d4[k] = s4[k];
    } else {
        for (int k = blockIdx.x * 256 + threadIdx.x; k < npix; k += gridDim.x * 256) d[k] = s[k];
    }
    if (blockIdx.x == 0 && threadIdx.x < 16) out_views[16 * r + threadIdx.x] = views[16 * (size_t)src + threadIdx.x];
    if (blockIdx.x == 0 && threadIdx.x < 3) out_rx[3 * r + threadIdx.x] = rx[3 * (size_t)src + threadIdx.x];
}

hipError_t launch_view_matrices(int N, const double* q, const float* rx, float* out, hipStream_t s)
{
    if (N == 0) return hipSuccess;
    hipLaunchKernelGGL(k_view_matrices, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, N, q[0], q[1], q[2], q[3],
                       rx, out);
    return hipGetLastError();
}

hipError_t launch_gather_batch(int R, int N, int npix, const int* order, int L, const uint32_t* train_state,
                               const float* spectra, const float* rx, const float* views, float* out_spec,
                               float* out_rx, float* out_views, hipStream_t s)
{
    if (R == 0) return hipSuccess;
    const int per = (npix & 3) == 0 ? npix / 4 : npix;
    const unsigned gx = (unsigned)std::max(1, std::min(64, (per + 255) / 256));
    hipLaunchKernelGGL(k_gather_batch, dim3(gx, (unsigned)R), dim3(256), 0, s, R, N, npix, order, L, train_state,
                       spectra, rx, views, out_spec, out_rx, out_views);
    return hipGetLastError();
}

}  // namespace gs
