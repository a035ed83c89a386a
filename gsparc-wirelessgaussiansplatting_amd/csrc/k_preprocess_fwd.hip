// Forward preprocess for R receivers at once (forward.cu:252-327 per receiver).
//
// A workgroup owns GS_PF_G consecutive Gaussians of one scene and one group of
// up to GS_RC receivers (blockIdx.y).  It first stages the group's receiver
// positions and view matrices and its Gaussians' 194 MLP weights in LDS: the
// weight sections are contiguous over consecutive Gaussians, so the staging
// loads are coalesced 16-byte loads and each weight is read from HBM once per
// group.  Wave w then takes receivers w*GS_PF_RW .. w*GS_PF_RW+GS_PF_RW-1 of
// the group, one Gaussian per lane: the emission MLP from the LDS copy, the
// projection, and the per-(Gaussian, receiver) records.  Receiver-independent
// work (cov3D, the log of the transmitter distance) is recomputed by each wave
// for its lane's Gaussian; the opacity's skip threshold (gs_alpha_thr: a
// search of a few floats around -log(255 o)) is found once per workgroup by
// wave 0 during the staging and shared through LDS.
#include "gs_device.h"
#include "gs_kernels.h"

namespace gs {

#define GS_PF_G 64                              // Gaussians per workgroup (one per lane)
#define GS_PF_RW 2                              // receivers per wave
#define GS_PF_WAVES (GS_RC / GS_PF_RW)
#define GS_PF_THREADS (GS_PF_WAVES * GS_WAVE)
#define GS_PF_ROW 49                            // float4 per staged Gaussian: W1 24 | b1 8 | W2 16 | b2

__global__ void __launch_bounds__(GS_PF_THREADS) k_preprocess_fwd(Dims d, PreFwdArgs a)
{
    const GroupMap grp = group_map(d, blockIdx.y);
    const SceneMap sc = scene_map(d, grp.scene);
    const int gb = blockIdx.x * GS_PF_G;                       // first Gaussian, within the scene
    if (gb >= (int)sc.P) return;
    const int tid = threadIdx.x, wave = tid / GS_WAVE, lane = tid % GS_WAVE;
    const int nG = min(GS_PF_G, (int)sc.P - gb);
    const int g = gb + lane;
    const bool valid = lane < nG;
    const int P = d.Ptot;                                      // MLP section stride
    const int gg0 = (int)sc.g0 + gb;                           // packed index of the first
    const int gg = gg0 + lane;
    const float* tx = a.tx + 3 * grp.scene;
    // the lane's Gaussian, loaded before the staging so that the loads overlap
    const int gc = gg0 + min(lane, nG - 1);
    const float mean[3] = {a.means[3 * gc], a.means[3 * gc + 1], a.means[3 * gc + 2]};
    const float sv[3] = {a.scales[3 * gc], a.scales[3 * gc + 1], a.scales[3 * gc + 2]};
    const float q[4] = {a.rots[4 * gc], a.rots[4 * gc + 1], a.rots[4 * gc + 2], a.rots[4 * gc + 3]};
    const float o = valid ? a.opac[gc] : 0.0f;

    // the group's receivers (zero beyond its nr: those slots are never used)
    __shared__ float s_rx[GS_RC * 3];
    __shared__ float s_view[GS_RC * 16];
    __shared__ float4 s_w[GS_PF_G * GS_PF_ROW];
    __shared__ float s_thr[GS_PF_G];
    for (int i = tid; i < GS_RC * 16; i += GS_PF_THREADS) {
        if (i < GS_RC * 3) s_rx[i] = i < 3 * grp.nr ? a.rx[3 * grp.r0 + i] : 0.0f;
        s_view[i] = i < 16 * grp.nr ? a.views[16 * grp.r0 + i] : 0.0f;
    }
    if (a.mlp) {
        // params = [W1 P*96 | b1 P*32 | W2 P*64 | b2 P*2] (gaussian.py:554-572);
        // the ABI checks the 16-byte alignment
        const float4* W1v = reinterpret_cast<const float4*>(a.mlp) + (size_t)gg0 * 24;
        const float4* b1v = reinterpret_cast<const float4*>(a.mlp + (size_t)P * 96) + (size_t)gg0 * 8;
        const float4* W2v = reinterpret_cast<const float4*>(a.mlp + (size_t)P * 128) + (size_t)gg0 * 16;
        const float2* b2v = reinterpret_cast<const float2*>(a.mlp + (size_t)P * 192) + gg0;
        // a thread's loads (3 + 1 + 2 + 1) are all issued before its LDS
        // stores, so that their latencies overlap.  The stores cover the
        // whole 64-row table unconditionally: rows past the block's nG
        // Gaussians get clamped duplicates, which no valid lane reads.
        constexpr int N1 = GS_PF_G * 24 / GS_PF_THREADS;
        constexpr int NB = GS_PF_G * 8 / GS_PF_THREADS;
        constexpr int N2 = GS_PF_G * 16 / GS_PF_THREADS;
        static_assert(N1 * GS_PF_THREADS == GS_PF_G * 24 && NB * GS_PF_THREADS == GS_PF_G * 8 &&
                          N2 * GS_PF_THREADS == GS_PF_G * 16 && GS_PF_G <= GS_PF_THREADS,
                      "the staging covers the table exactly");
        float4 w1[N1], bb[NB], w2[N2];
        float2 b2 = make_float2(0.0f, 0.0f);
#pragma unroll
        for (int u = 0; u < N1; ++u) w1[u] = W1v[min(tid + u * GS_PF_THREADS, nG * 24 - 1)];
#pragma unroll
        for (int u = 0; u < NB; ++u) bb[u] = b1v[min(tid + u * GS_PF_THREADS, nG * 8 - 1)];
#pragma unroll
        for (int u = 0; u < N2; ++u) w2[u] = W2v[min(tid + u * GS_PF_THREADS, nG * 16 - 1)];
        if (tid < GS_PF_G) b2 = b2v[min(tid, nG - 1)];
#pragma unroll
        for (int u = 0; u < N1; ++u) {
            const int i = tid + u * GS_PF_THREADS;
            s_w[(i / 24) * GS_PF_ROW + i % 24] = w1[u];
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int i = tid + u * GS_PF_THREADS;
            s_w[(i / 8) * GS_PF_ROW + 24 + i % 8] = bb[u];
        }
#pragma unroll
        for (int u = 0; u < N2; ++u) {
            const int i = tid + u * GS_PF_THREADS;
            s_w[(i / 16) * GS_PF_ROW + 32 + i % 16] = w2[u];
        }
        if (tid < GS_PF_G) s_w[tid * GS_PF_ROW + 48] = make_float4(b2.x, b2.y, 0.0f, 0.0f);
    }
    // alpha = o*G >= 1/255 exactly when power >= thr (gs_alpha_thr)
    if (wave == 0) s_thr[lane] = gs_alpha_thr(o);
    __syncthreads();
    const int c0 = wave * GS_PF_RW;
    if (c0 >= grp.nr) return;                                  // no barrier below
    const int c1 = min(grp.nr, c0 + GS_PF_RW);

    float c3[6] = {0, 0, 0, 0, 0, 0}, ld = 0.0f;
    if (valid) {
        cov3d(sv, a.mod, q, c3);
        if (wave == 0 && grp.r0 == (int)sc.r0)   // the scene's first group, once
#pragma unroll
            for (int k = 0; k < 6; ++k) a.cov3[6 * gg + k] = c3[k];
        const float df[3] = {tx[0] - mean[0], tx[1] - mean[1], tx[2] - mean[2]};
        const float dist = sqrtf(df[0] * df[0] + df[1] * df[1] + df[2] * df[2]);
        ld = gs_logf(dist + 1e-6f);
    }
    const float thr = s_thr[lane];

    // ---- emission MLP 3->32->2 for this wave's receivers (forward.cu:17-75):
    // four hidden units per step, dot products as fused multiply-adds in
    // index order (forward.cu:55-70)
    float z0[GS_PF_RW], z1[GS_PF_RW];
    if (a.mlp && valid) {
        const float4* wr = s_w + lane * GS_PF_ROW;
        float rxr[GS_PF_RW][3];
        const float4 b2 = wr[48];
#pragma unroll
        for (int k = 0; k < GS_PF_RW; ++k) {
            z0[k] = b2.x; z1[k] = b2.y;
            rxr[k][0] = s_rx[3 * (c0 + k)]; rxr[k][1] = s_rx[3 * (c0 + k) + 1]; rxr[k][2] = s_rx[3 * (c0 + k) + 2];
        }
        for (int j4 = 0; j4 < 8; ++j4) {
            const float4 wa = wr[3 * j4], wb = wr[3 * j4 + 1], wc = wr[3 * j4 + 2];
            const float4 bq = wr[24 + j4], va = wr[32 + j4], vb = wr[40 + j4];
            const float w[12] = {wa.x, wa.y, wa.z, wa.w, wb.x, wb.y, wb.z, wb.w, wc.x, wc.y, wc.z, wc.w};
            const float bu[4] = {bq.x, bq.y, bq.z, bq.w};
            const float v0u[4] = {va.x, va.y, va.z, va.w}, v1u[4] = {vb.x, vb.y, vb.z, vb.w};
#pragma unroll
            for (int u = 0; u < 4; ++u) {
#pragma unroll
                for (int k = 0; k < GS_PF_RW; ++k) {
                    float sm = fmaf(w[3 * u], rxr[k][0], bu[u]);
                    sm = fmaf(w[3 * u + 1], rxr[k][1], sm);
                    sm = fmaf(w[3 * u + 2], rxr[k][2], sm);
                    // leaky ReLU, slope 0.01 < 1: max(s, 0.01 s)
                    const float h = fmaxf(sm, 0.01f * sm);
                    z0[k] = fmaf(v0u[u], h, z0[k]);
                    z1[k] = fmaf(v1u[u], h, z1[k]);
                }
            }
        }
    }
    if (!valid) return;
#pragma unroll
    for (int k = 0; k < GS_PF_RW; ++k) {
        const int c = c0 + k;
        if (c >= c1) break;
        const int r = grp.r0 + c;
        const float* view = s_view + 16 * c;
        const size_t ri = (size_t)rx_map(d, r).roff + g;
        int x0 = 0, y0 = 0, x1 = 0, y1 = 0;
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
                    if (a.plen) a.plen[ri] = gs_path_length(mean[0], mean[1], mean[2], tx, s_rx + 3 * c);
                    if (a.mlp) {
                        const float s0 = 1.0f / (1.0f + gs_expf(-(z0[k] - ld)));
                        const float s1 = 1.0f / (1.0f + gs_expf(-(z1[k] - ld)));
                        a.sig[ri] = make_float2(s0, s1);
                    }
                }
            }
        }
        a.radii[ri] = radius;
    }
}

hipError_t launch_preprocess_fwd(const Dims& d, const PreFwdArgs& a, hipStream_t s)
{
    const int NB = (d.P + GS_PF_G - 1) / GS_PF_G;
    if (NB == 0 || d.G == 0) return hipSuccess;
    hipLaunchKernelGGL(k_preprocess_fwd, dim3(NB, d.G), dim3(GS_PF_THREADS), 0, s, d, a);
    return hipGetLastError();
}

}  // namespace gs
