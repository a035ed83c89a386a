// Forward preprocess for R receivers at once (forward.cu:252-327 per receiver).
//
// One thread per Gaussian.  Receiver-independent work (cov3D, the log of the
// transmitter distance, opacity threshold) is done once; the emission MLP is
// evaluated for groups of up to GS_RC receivers of one scene (one group per
// blockIdx.y) so that each Gaussian's 194 weights are streamed once per group,
// not per receiver.  blockIdx.x * 256 + threadIdx.x is the Gaussian's index
// within the group's scene.
#include "gs_device.h"
#include "gs_kernels.h"

namespace gs {

__global__ void __launch_bounds__(GS_PRE_THREADS)
k_preprocess_fwd(Dims d, PreFwdArgs a)
{
    const GroupMap grp = group_map(d, blockIdx.y);
    const SceneMap sc = scene_map(d, grp.scene);
    const int g = blockIdx.x * GS_PRE_THREADS + threadIdx.x;   // within the scene
    if (blockIdx.x * GS_PRE_THREADS >= (int)sc.P) return;
    const bool valid = g < (int)sc.P;
    const int P = d.Ptot;                                      // MLP section stride
    const int gg = (int)sc.g0 + g;                             // packed index
    const float* tx = a.tx + 3 * grp.scene;

    float mean[3] = {0, 0, 0}, c3[6] = {0, 0, 0, 0, 0, 0}, o = 0.0f, ld = 0.0f;
    if (valid) {
        mean[0] = a.means[3 * gg]; mean[1] = a.means[3 * gg + 1]; mean[2] = a.means[3 * gg + 2];
        const float s[3] = {a.scales[3 * gg], a.scales[3 * gg + 1], a.scales[3 * gg + 2]};
        const float q[4] = {a.rots[4 * gg], a.rots[4 * gg + 1], a.rots[4 * gg + 2], a.rots[4 * gg + 3]};
        cov3d(s, a.mod, q, c3);
        if (grp.r0 == (int)sc.r0)   // the scene's first group
#pragma unroll
            for (int k = 0; k < 6; ++k) a.cov3[6 * gg + k] = c3[k];
        o = a.opac[gg];
        const float df[3] = {tx[0] - mean[0], tx[1] - mean[1], tx[2] - mean[2]};
        const float dist = sqrtf(df[0] * df[0] + df[1] * df[1] + df[2] * df[2]);
        ld = gs_logf(dist + 1e-6f);
    }
    // alpha = o*G < 1/255 for every power < thr (see gs_device.h)
    const float thr = -gs_logf(255.0f * o) - 1e-3f;

    {
        const int r0 = grp.r0;
        const int nr = grp.nr;
        // ---- emission MLP 3->32->2 for receivers r0..r0+nr (forward.cu:17-75)
        float z0[GS_RC], z1[GS_RC];
        if (a.mlp && valid) {
            // the Gaussian's weights as 16-byte loads (the ABI checks the
            // alignment), four hidden units per step, summed in unit order
            const float4* W1v = reinterpret_cast<const float4*>(a.mlp + (size_t)gg * 96);
            const float4* b1v = reinterpret_cast<const float4*>(a.mlp + (size_t)P * 96 + (size_t)gg * 32);
            const float4* W2v = reinterpret_cast<const float4*>(a.mlp + (size_t)P * 128 + (size_t)gg * 64);
            const float2 b2 = *reinterpret_cast<const float2*>(a.mlp + (size_t)P * 192 + (size_t)gg * 2);
#pragma unroll
            for (int c = 0; c < GS_RC; ++c) { z0[c] = b2.x; z1[c] = b2.y; }
            for (int j4 = 0; j4 < 8; ++j4) {
                const float4 wa = W1v[3 * j4], wb = W1v[3 * j4 + 1], wc = W1v[3 * j4 + 2];
                const float4 bq = b1v[j4], va = W2v[j4], vb = W2v[8 + j4];
                const float w[12] = {wa.x, wa.y, wa.z, wa.w, wb.x, wb.y, wb.z, wb.w, wc.x, wc.y, wc.z, wc.w};
                const float bu[4] = {bq.x, bq.y, bq.z, bq.w};
                const float v0u[4] = {va.x, va.y, va.z, va.w}, v1u[4] = {vb.x, vb.y, vb.z, vb.w};
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const float w0 = w[3 * u], w1 = w[3 * u + 1], w2 = w[3 * u + 2], bb = bu[u];
                    const float v0 = v0u[u], v1 = v1u[u];
#pragma unroll
                    for (int c = 0; c < GS_RC; ++c) {
                        if (c < nr) {
                            const float* rx = a.rx + 3 * (r0 + c);
                            float s = bb;
                            s = s + w0 * rx[0];
                            s = s + w1 * rx[1];
                            s = s + w2 * rx[2];
                            const float h = (s > 0.0f) ? s : 0.01f * s;
                            z0[c] = z0[c] + v0 * h;
                            z1[c] = z1[c] + v1 * h;
                        }
                    }
                }
            }
        }
        for (int c = 0; c < nr; ++c) {
            const int r = r0 + c;
            const float* view = a.views + 16 * r;
            const size_t ri = (size_t)rx_map(d, r).roff + g;
            int x0 = 0, y0 = 0, x1 = 0, y1 = 0;
            if (valid) {
                int radius = 0;
                float t[3];
                view_xform(mean[0], mean[1], mean[2], view, t);
                const float rr = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
                if (!(rr <= 0.04f)) {                              // auxiliary.h:143
                    GeoOut go;
                    const float rd = sqrtf(rr);
                    project(t, rd, d.W, d.H, go);
                    float M2[2][3], A, B, C;
                    cov2d(go.J, view, c3, M2, A, B, C);
                    const float det = A * C - B * B;
                    if (det != 0.0f) {                             // forward.cu:296
                        const float det_inv = 1.0f / det;
                        const float mid = 0.5f * (A + C);
                        const float disc = sqrtf(fmaxf(0.1f, mid * mid - det));
                        const float l1 = mid + disc;
                        radius = (int)ceilf(3.0f * sqrtf(l1));
                        get_rect(go.px, go.py, radius, d.gx, d.gy, x0, y0, x1, y1);
                        if ((x1 - x0) * (y1 - y0) == 0) {          // forward.cu:311
                            radius = 0;
                            x0 = y0 = x1 = y1 = 0;
                        } else {
                            const float qa = C * det_inv, qb = -B * det_inv, qc = A * det_inv;
                            a.splat[2 * ri] = make_float4(go.px, go.py, -0.5f * qa, -qb);
                            a.splat[2 * ri + 1] = make_float4(-0.5f * qc, o, thr, 0.0f);
                            a.depth[ri] = rd;
                            if (a.rect)
                                a.rect[ri] = (uint32_t)x0 | ((uint32_t)y0 << 8) | ((uint32_t)x1 << 16) |
                                             ((uint32_t)y1 << 24);
                            if (a.plen) a.plen[ri] = gs_path_length(mean[0], mean[1], mean[2], tx, a.rx + 3 * r);
                            if (a.mlp) {
                                const float s0 = 1.0f / (1.0f + gs_expf(-(z0[c] - ld)));
                                const float s1 = 1.0f / (1.0f + gs_expf(-(z1[c] - ld)));
                                a.sig[ri] = make_float2(s0, s1);
                            }
                        }
                    }
                }
                a.radii[ri] = radius;
            }
        }
    }
}

hipError_t launch_preprocess_fwd(const Dims& d, const PreFwdArgs& a, hipStream_t s)
{
    const int NB = (d.P + GS_PRE_THREADS - 1) / GS_PRE_THREADS;
    if (NB == 0 || d.G == 0) return hipSuccess;
    hipLaunchKernelGGL(k_preprocess_fwd, dim3(NB, d.G), dim3(GS_PRE_THREADS), 0, s, d, a);
    return hipGetLastError();
}

}  // namespace gs
