// Preprocess backward for R receivers (computeCov2DCUDA backward.cu:124-303,
// preprocessCUDA backward.cu:378-437, computeCov3D backward.cu:308-371,
// computeSignalFromMLP backward backward.cu:9-118).
//
// Two kernels, no atomics (summation order is fixed): k_pre_bwd_geo runs the
// geometry adjoint per (Gaussian, chunk of receivers) and parks partial sums
// and the MLP output adjoints in the grad workspace; k_pre_bwd_param forms the
// 194 MLP weight gradients with one thread per (Gaussian, hidden unit), sums
// the partials and runs the cov3D adjoint once on the receiver sum.
//
// Modes: GRAD_EXACT is the true adjoint of the forward; GRAD_REFERENCE
// reproduces the reference's backward literally (halved x-row of the
// Jacobian, 1/(det^2 + 1e-7), 1/(d+eps)^2 in the MLP distance gradient,
// scale modifier dropped from dL/dscale).
#include "gs_device.h"
#include "gs_kernels.h"

namespace gs {

// The geometry adjoint only produces gradients (held to 1e-4 of the largest
// magnitude), so its quotients and square roots use the hardware reciprocal
// and square root (1 ulp) instead of the correctly rounded sequences.
__device__ __forceinline__ float rcp_g(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float sqrt_g(float x) { return __builtin_amdgcn_sqrtf(x); }

// The Jacobian of the equirectangular projection (as project() forms it)
// with those operations.
__device__ __forceinline__ void jacobian_g(const float* t, int W, int H, float J[2][3])
{
    const float rxz2 = t[0] * t[0] + t[2] * t[2];
    const float rxz2_inv = rcp_g(rxz2 + 1e-7f);
    const float rxz = sqrt_g(rxz2);
    const float rxz_inv = rcp_g(rxz + 1e-7f);
    const float r2 = rxz2 + t[1] * t[1];
    const float r2_inv = rcp_g(r2 + 1e-7f);
    const float Wd = (float)W * 0.5f * GS_ONE_OVER_PI_F;
    const float Hd = (float)H * GS_ONE_OVER_PI_F;
    J[0][0] = Wd * t[2] * rxz2_inv;
    J[0][1] = 0.0f;
    J[0][2] = -Wd * t[0] * rxz2_inv;
    J[1][0] = -Hd * t[0] * t[1] * rxz_inv * r2_inv;
    J[1][1] = Hd * rxz * r2_inv;
    J[1][2] = -Hd * t[2] * t[1] * rxz_inv * r2_inv;
}

#define GS_BWD_THREADS 64
#define GS_PART 12   // per (receiver chunk, Gaussian): dmean[3], dcov3[6], dopacity, db2[2]

template <int MODE>
__device__ __forceinline__ void geo_bwd(const float* mean, const float* c3, const float* view, int W, int H,
                                        const float* g, float* dmean, float* dc)
{
    float t[3];
    view_xform(mean[0], mean[1], mean[2], view, t);
    const float rr = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
    const float rd = sqrt_g(rr);
    const float dpx = g[0], dpy = g[1], gqa = g[2], gqb = g[3], gqc = g[4];
    float Jc[2][3];
    jacobian_g(t, W, H, Jc);
    if (MODE == GRAD_REFERENCE) {   // backward.cu:172-173
        Jc[0][0] *= 0.5f;
        Jc[0][2] *= 0.5f;
    }
    float m[2][3], A, B, C;
    cov2d(Jc, view, c3, m, A, B, C);
    const float det = A * C - B * B;
    float dA, dB, dC;
    if (MODE == GRAD_REFERENCE) {   // backward.cu:212-224
        const float d2i = rcp_g(det * det + 1e-7f);
        const float gy = 0.5f * gqb;
        dA = d2i * (-C * C * gqa + 2.0f * B * C * gy + (det - A * C) * gqc);
        dC = d2i * (-A * A * gqc + 2.0f * A * B * gy + (det - A * C) * gqa);
        dB = d2i * 2.0f * (B * C * gqa - (det + 2.0f * B * B) * gy + A * B * gqc);
    } else {
        const float d2i = rcp_g(det * det);
        dA = d2i * (-C * C * gqa + B * C * gqb - B * B * gqc);
        dB = d2i * (2.0f * B * C * gqa - (det + 2.0f * B * B) * gqb + 2.0f * A * B * gqc);
        dC = d2i * (-B * B * gqa + A * B * gqb - A * A * gqc);
    }
    const int ij[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
#pragma unroll
    for (int e = 0; e < 6; ++e) {
        const int i = ij[e][0], j = ij[e][1];
        if (i == j)
            dc[e] += m[0][i] * m[0][i] * dA + m[0][i] * m[1][i] * dB + m[1][i] * m[1][i] * dC;
        else
            dc[e] += 2.0f * m[0][i] * m[0][j] * dA + (m[0][i] * m[1][j] + m[0][j] * m[1][i]) * dB +
                     2.0f * m[1][i] * m[1][j] * dC;
    }
    const float S[3][3] = {{c3[0], c3[1], c3[2]}, {c3[1], c3[3], c3[4]}, {c3[2], c3[4], c3[5]}};
    float Sm0[3], Sm1[3], dm0[3], dm1[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        Sm0[i] = S[i][0] * m[0][0] + S[i][1] * m[0][1] + S[i][2] * m[0][2];
        Sm1[i] = S[i][0] * m[1][0] + S[i][1] * m[1][1] + S[i][2] * m[1][2];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        dm0[i] = 2.0f * dA * Sm0[i] + dB * Sm1[i];
        dm1[i] = 2.0f * dC * Sm1[i] + dB * Sm0[i];
    }
    float dJ[2][3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        dJ[0][q] = dm0[0] * view[0 * 4 + q] + dm0[1] * view[1 * 4 + q] + dm0[2] * view[2 * 4 + q];
        dJ[1][q] = dm1[0] * view[0 * 4 + q] + dm1[1] * view[1 * 4 + q] + dm1[2] * view[2 * 4 + q];
    }
    const float Wd = (float)W * 0.5f * GS_ONE_OVER_PI_F;
    const float Hd = (float)H * GS_ONE_OVER_PI_F;
    const float tx = t[0], ty = t[1], tz = t[2];
    float dt[3] = {0.0f, 0.0f, 0.0f};
    if (MODE == GRAD_REFERENCE) {
        // backward.cu:152-177, 249-292
        const float txtx = tx * tx, tyty = ty * ty, tztz = tz * tz, txtytz = tx * ty * tz;
        const float rxz2 = txtx + tztz;
        const float rxz2_inv = rcp_g(rxz2 + 1e-7f);
        const float rxz4_inv = rxz2_inv * rxz2_inv;
        const float rxz = sqrt_g(rxz2);
        const float rxz_inv = rcp_g(rxz + 1e-7f);
        const float r2 = rxz2 + tyty;
        const float r2_inv = rcp_g(r2 + 1e-7f);
        const float r4_inv = r2_inv * r2_inv;
        const float rxz_r4_inv = rxz_inv * r4_inv;
        const float rxz3_r4_inv = rxz2_inv * rxz_r4_inv;
        const float temp1 = Hd * (tyty - rxz2) * rxz_r4_inv;
        const float temp2 = Hd * txtytz * (r2 + 2.0f * rxz2) * rxz3_r4_inv;
        const float temp3 = Wd * (txtx - tztz) * rxz4_inv;
        const float temp4 = Wd * 2.0f * tx * tz * rxz4_inv;
        const float temp5 = Hd * ty * rxz3_r4_inv;
        dt[0] = -dJ[0][0] * temp4 + dJ[0][2] * temp3 + dJ[1][0] * temp5 * (2.0f * txtx * rxz2 - tztz * r2) +
                dJ[1][1] * tx * temp1 + dJ[1][2] * temp2;
        dt[1] = dJ[1][0] * tx * temp1 - dJ[1][1] * Hd * 2.0f * rxz * ty * r4_inv + dJ[1][2] * tz * temp1;
        dt[2] = dJ[0][0] * temp3 + dJ[0][2] * temp4 + dJ[1][0] * temp2 + dJ[1][1] * tz * temp1 +
                dJ[1][2] * temp5 * (2.0f * tztz * rxz2 - txtx * r2);
        dt[0] += dpx * Jc[0][0] + dpy * Jc[1][0];
        dt[1] += dpx * Jc[0][1] + dpy * Jc[1][1];
        dt[2] += dpx * Jc[0][2] + dpy * Jc[1][2];
    } else {
        const float eps = 1e-7f;
        const float rxz2 = tx * tx + tz * tz;
        const float ia = rcp_g(rxz2 + eps);
        const float rho = sqrt_g(rxz2);
        const float ib = rcp_g(rho + eps);
        const float r2 = rxz2 + ty * ty;
        const float ic = rcp_g(r2 + eps);
        const float g0x = dJ[0][0], g0z = dJ[0][2], h0 = dJ[1][0], h1 = dJ[1][1], h2 = dJ[1][2];
        dt[0] += g0x * Wd * tz * (-2.0f * tx * ia * ia) + g0z * (-Wd) * (ia - 2.0f * tx * tx * ia * ia);
        dt[2] += g0x * Wd * (ia - 2.0f * tz * tz * ia * ia) + g0z * (-Wd) * tx * (-2.0f * tz * ia * ia);
        const float Kc = ib * ic;
        const float irho = rcp_g(rho);
        const float drho_x = rho > 0.0f ? tx * irho : 0.0f, drho_z = rho > 0.0f ? tz * irho : 0.0f;
        const float dK0 = -ic * ib * ib * drho_x - 2.0f * ib * ic * ic * tx;
        const float dK1 = -2.0f * ib * ic * ic * ty;
        const float dK2 = -ic * ib * ib * drho_z - 2.0f * ib * ic * ic * tz;
        dt[0] += h0 * (-Hd) * (ty * Kc + tx * ty * dK0);
        dt[1] += h0 * (-Hd) * (tx * Kc + tx * ty * dK1);
        dt[2] += h0 * (-Hd) * (tx * ty * dK2);
        dt[0] += h2 * (-Hd) * (tz * ty * dK0);
        dt[1] += h2 * (-Hd) * (tz * Kc + tz * ty * dK1);
        dt[2] += h2 * (-Hd) * (ty * Kc + tz * ty * dK2);
        dt[0] += h1 * Hd * (drho_x * ic - 2.0f * rho * tx * ic * ic);
        dt[1] += h1 * Hd * (-2.0f * rho * ty * ic * ic);
        dt[2] += h1 * Hd * (drho_z * ic - 2.0f * rho * tz * ic * ic);
        const float Wdp = (float)W * 0.5f * GS_ONE_OVER_PI_F;
        const float Hdp = (float)H * 0.5f * GS_TWO_OVER_PI_F;
        if (rxz2 > 0.0f) {
            const float irxz2 = rcp_g(rxz2);
            dt[0] += dpx * Wdp * tz * irxz2;
            dt[2] += dpx * Wdp * (-tx) * irxz2;
        }
        const float re = rd + eps;
        const float ire = rcp_g(re);
        const float u = ty * ire;
        const float dlat_du = rcp_g(sqrt_g(1.0f - u * u));
        const float k2 = ty * rcp_g(rd * re * re);
        dt[0] += dpy * Hdp * dlat_du * (-k2 * tx);
        dt[1] += dpy * Hdp * dlat_du * (ire - k2 * ty);
        dt[2] += dpy * Hdp * dlat_du * (-k2 * tz);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) dmean[k] += view[k * 4 + 0] * dt[0] + view[k * 4 + 1] * dt[1] + view[k * 4 + 2] * dt[2];
}

// Kernel 1: one thread per (Gaussian, receiver group of its scene).  Runs the
// geometry adjoint per receiver (with the MLP outputs and cov3D the forward
// stored) and writes the group's partial sums (dmean[3], dcov3[6], dopacity,
// db2[2]) plus dz[r] = dL/d(MLP output) for the weight-gradient kernel.
template <int MODE>
__global__ void __launch_bounds__(GS_BWD_THREADS) k_pre_bwd_geo(Dims d, PreBwdArgs a)
{
    const GroupMap grp = group_map(d, blockIdx.y);
    const SceneMap sc = scene_map(d, grp.scene);
    const int g = blockIdx.x * GS_BWD_THREADS + threadIdx.x;   // within the scene
    const int r0 = grp.r0;
    const int nr = grp.nr;
    // the group's view matrices, receiver positions and record offsets, read
    // once per block: inside the receiver loop the only vector-memory loads
    // are then the (prefetched) per-pair records
    __shared__ float s_view[GS_RC * 16];
    __shared__ float s_rx[GS_RC * 3];
    __shared__ uint32_t s_roff[GS_RC];
    if (blockIdx.x * GS_BWD_THREADS >= (int)sc.P) return;
    for (int i = threadIdx.x; i < GS_RC * 16; i += GS_BWD_THREADS) {
        s_view[i] = i < 16 * nr ? a.views[16 * r0 + i] : 0.0f;
        if (i < GS_RC * 3) s_rx[i] = i < 3 * nr ? a.rx[3 * r0 + i] : 0.0f;
        if (i < GS_RC) s_roff[i] = i < nr ? rx_map(d, r0 + i).roff : 0u;
    }
    __syncthreads();
    if (g >= (int)sc.P) return;   // no block-level synchronisation below
    const int gg = (int)sc.g0 + g;
    const float mean[3] = {a.means[3 * gg], a.means[3 * gg + 1], a.means[3 * gg + 2]};
    // cov3D as the forward stored it
    float c3[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) c3[k] = a.cov3[6 * gg + k];
    const float* tx = a.tx + 3 * grp.scene;
    const float diff[3] = {tx[0] - mean[0], tx[1] - mean[1], tx[2] - mean[2]};
    const float dist = sqrt_g(diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]);
    const float de = dist + 1e-6f;
    const float iden = rcp_g((MODE == GRAD_REFERENCE) ? de * de : dist * de);
    const float idist = rcp_g(dist);

    float dmean[3] = {0, 0, 0}, dc[6] = {0, 0, 0, 0, 0, 0}, dop = 0.0f, db2_0 = 0.0f, db2_1 = 0.0f;
    // The records of receiver c + 1 are loaded (unconditionally: a culled
    // pair's record is read and discarded) while receiver c is computed.
    const float4* g8v = reinterpret_cast<const float4*>(a.g8);
    // the conic (ha, nb, hc) of each record, from its splat record
    const float* spf = reinterpret_cast<const float*>(a.splat);
    size_t ri_n = (size_t)s_roff[0] + g;
    int rad_n = a.radii[ri_n];
    float4 ga_n = g8v[2 * ri_n], gb_n = g8v[2 * ri_n + 1];
    float2 cq_n = *reinterpret_cast<const float2*>(spf + 8 * ri_n + 2);
    float ch_n = spf[8 * ri_n + 4];
    float2 sg_n = a.mlp ? a.sig[ri_n] : make_float2(0.0f, 0.0f);
#pragma unroll
    for (int c = 0; c < GS_RC; ++c) {
        if (c >= nr) break;
        const int r = r0 + c;
        const size_t ri = ri_n;
        const int rad = rad_n;
        const float4 ga = ga_n, gb = gb_n;
        const float2 cq = cq_n;
        const float ch = ch_n;
        const float2 sg = sg_n;
        if (c + 1 < nr) {
            ri_n = (size_t)s_roff[c + 1] + g;
            rad_n = a.radii[ri_n];
            ga_n = g8v[2 * ri_n];
            gb_n = g8v[2 * ri_n + 1];
            cq_n = *reinterpret_cast<const float2*>(spf + 8 * ri_n + 2);
            ch_n = spf[8 * ri_n + 4];
            if (a.mlp) sg_n = a.sig[ri_n];
        }
        float2 dz = make_float2(0.0f, 0.0f);
        if (rad > 0) {
            // dL/dmean2D = D (conic . d) with the conic (qa, qb; qb, qc) =
            // (-2 ha, -nb; -nb, -2 hc) (backward.cu:600-605), and dL/dconic
            // = (-1/2 D dx^2, -D dx dy, -1/2 D dy^2) (backward.cu:608-610; the
            // off-diagonal's 1/2 is applied below)
            const float ha = cq.x, nb = cq.y, hc = ch;
            const float gv[8] = {fmaf(nb, ga.y, (2.0f * ha) * ga.x), fmaf(nb, ga.x, (2.0f * hc) * ga.y),
                                 -0.5f * ga.z, -ga.w, -0.5f * gb.x, gb.y, gb.z, gb.w};
            geo_bwd<MODE>(mean, c3, s_view + 16 * c, d.W, d.H, gv, dmean, dc);
            dop += gv[5];
            if (a.gL) {
                // CSI phase: L = |mu - tx| + |mu - rx|
                const float gl = a.gL[ri];
                const float* rxp = s_rx + 3 * c;
                const float b[3] = {mean[0] - rxp[0], mean[1] - rxp[1], mean[2] - rxp[2]};
                const float idb = rcp_g(sqrt_g(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]));
#pragma unroll
                for (int k = 0; k < 3; ++k) dmean[k] += gl * (b[k] * idb - diff[k] * idist);
            }
            if (a.dmeans2d) {
                a.dmeans2d[3 * ri + 0] = gv[0] * (0.5f * (float)d.W);
                a.dmeans2d[3 * ri + 1] = gv[1] * (0.5f * (float)d.H);
                a.dmeans2d[3 * ri + 2] = 0.0f;
            }
            if (a.mlp) {
                // MLP outputs as the forward computed them
                const float s0 = sg.x, s1 = sg.y;
                dz.x = gv[6] * s0 * (1.0f - s0);
                dz.y = gv[7] * s1 * (1.0f - s1);
                const float dld = -(dz.x + dz.y);
#pragma unroll
                for (int k = 0; k < 3; ++k) dmean[k] += -dld * diff[k] * iden;
                db2_0 += dz.x;
                db2_1 += dz.y;
            } else if (a.dsig) {
                a.dsig[2 * ri] = gv[6];
                a.dsig[2 * ri + 1] = gv[7];
            }
        } else {
            if (a.dmeans2d) {
                a.dmeans2d[3 * ri + 0] = 0.0f;
                a.dmeans2d[3 * ri + 1] = 0.0f;
                a.dmeans2d[3 * ri + 2] = 0.0f;
            }
            if (!a.mlp && a.dsig) {
                a.dsig[2 * ri] = 0.0f;
                a.dsig[2 * ri + 1] = 0.0f;
            }
        }
        if (a.mlp) a.dz[ri] = dz;
    }
    float* pt = a.part + ((size_t)grp.part + g) * GS_PART;
    pt[0] = dmean[0]; pt[1] = dmean[1]; pt[2] = dmean[2];
#pragma unroll
    for (int k = 0; k < 6; ++k) pt[3 + k] = dc[k];
    pt[9] = dop; pt[10] = db2_0; pt[11] = db2_1;
}

// Receiver-summed partials of Gaussian gl of scene sc (packed index gg):
// dmean, dopacity and db2 go out as they are; the cov3D adjoint
// (backward.cu:308-371) is linear in dL/dcov3D and so runs once, on the sum.
__device__ __forceinline__ void pre_bwd_finish(const Dims& d, const PreBwdArgs& a, const SceneMap& sc, int gl, int gg)
{
    const int P = d.Ptot;   // MLP section stride
    float acc[GS_PART];
#pragma unroll
    for (int k = 0; k < GS_PART; ++k) acc[k] = 0.0f;
    for (int y = (int)sc.group0; y < (int)(sc.group0 + sc.ngroups); ++y) {
        const float* pt = a.part + ((size_t)group_map(d, y).part + gl) * GS_PART;
#pragma unroll
        for (int k = 0; k < GS_PART; ++k) acc[k] += pt[k];
    }
    if (a.mlp && a.dmlp) {
        a.dmlp[(size_t)P * 192 + (size_t)gg * 2] = acc[10];
        a.dmlp[(size_t)P * 192 + (size_t)gg * 2 + 1] = acc[11];
    }
    const float* dc = acc + 3;
    const float sv[3] = {a.scales[3 * gg], a.scales[3 * gg + 1], a.scales[3 * gg + 2]};
    const float q[4] = {a.rots[4 * gg], a.rots[4 * gg + 1], a.rots[4 * gg + 2], a.rots[4 * gg + 3]};
    {
        const float r = q[0], x = q[1], y = q[2], z = q[3];
        const float Rq[3][3] = {
            {1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y - r * z), 2.0f * (x * z + r * y)},
            {2.0f * (x * y + r * z), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z - r * x)},
            {2.0f * (x * z - r * y), 2.0f * (y * z + r * x), 1.0f - 2.0f * (x * x + y * y)}};
        const float se[3] = {a.mod * sv[0], a.mod * sv[1], a.mod * sv[2]};
        const float Dg[3] = {se[0] * se[0], se[1] * se[1], se[2] * se[2]};
        const float G[3][3] = {{dc[0], 0.5f * dc[1], 0.5f * dc[2]},
                               {0.5f * dc[1], dc[3], 0.5f * dc[4]},
                               {0.5f * dc[2], 0.5f * dc[4], dc[5]}};
        float GR[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int k = 0; k < 3; ++k)
                GR[i][k] = 2.0f * (G[i][0] * Rq[0][k] + G[i][1] * Rq[1][k] + G[i][2] * Rq[2][k]) * Dg[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float dD = 0.0f;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int jj = 0; jj < 3; ++jj) dD += Rq[i][k] * G[i][jj] * Rq[jj][k];
            // dD/ds = 2 mod^2 s; the reference drops one factor mod (backward.cu:353-355)
            const float f = (a.mode == GRAD_REFERENCE) ? 2.0f * se[k] : 2.0f * se[k] * a.mod;
            a.dscales[3 * gg + k] = dD * f;
        }
        const float dRr[3][3] = {{0, -2 * z, 2 * y}, {2 * z, 0, -2 * x}, {-2 * y, 2 * x, 0}};
        const float dRx[3][3] = {{0, 2 * y, 2 * z}, {2 * y, -4 * x, -2 * r}, {2 * z, 2 * r, -4 * x}};
        const float dRy[3][3] = {{-4 * y, 2 * x, 2 * r}, {2 * x, 0, 2 * z}, {-2 * r, 2 * z, -4 * y}};
        const float dRz[3][3] = {{-4 * z, -2 * r, 2 * x}, {2 * r, -4 * z, 2 * y}, {2 * x, 2 * y, 0}};
        float dq[4] = {0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int jj = 0; jj < 3; ++jj) {
                dq[0] += GR[i][jj] * dRr[i][jj];
                dq[1] += GR[i][jj] * dRx[i][jj];
                dq[2] += GR[i][jj] * dRy[i][jj];
                dq[3] += GR[i][jj] * dRz[i][jj];
            }
#pragma unroll
        for (int k = 0; k < 4; ++k) a.drots[4 * gg + k] = dq[k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) a.dmeans[3 * gg + k] = acc[k];
    a.dopac[gg] = acc[9];
    if (a.dcov3)
#pragma unroll
        for (int k = 0; k < 6; ++k) a.dcov3[6 * gg + k] = dc[k];
}

// Kernel 2: the 194 MLP weight gradients, summed over the receivers of the
// scene (backward.cu:95-115).  A workgroup owns GS_PB_G consecutive Gaussians
// of one scene (blockIdx.y), one per lane; wave w takes hidden units
// 4w .. 4w+3.  Per chunk of 64 receivers the workgroup stages their
// positions and the block's dL/d(MLP output) in LDS (coalesced loads, each
// value read from memory once for all 8 waves).  Wave 0 then sums the group
// partials of its lane's Gaussian and runs the cov3D adjoint.
#define GS_PB_G 64
#define GS_PB_THREADS 512    // 8 waves x 4 hidden units
__global__ void __launch_bounds__(GS_PB_THREADS) k_pre_bwd_param(Dims d, PreBwdArgs a)
{
    const SceneMap sc = scene_map(d, blockIdx.y);
    const int gb = blockIdx.x * GS_PB_G;
    if (gb >= (int)sc.P) return;
    const int tid = threadIdx.x, wave = tid / GS_WAVE, lane = tid % GS_WAVE;
    const int gl = gb + lane;                 // within the scene
    const bool valid = gl < (int)sc.P;
    const int gg = (int)sc.g0 + gl;           // packed
    const int P = d.Ptot;                     // MLP section stride
    if (a.mlp && a.dmlp) {
        __shared__ float s_rx[GS_WAVE * 3];
        __shared__ uint32_t s_roff[GS_WAVE];
        // dL/d(MLP output) of the block's Gaussians for a chunk of 64
        // receivers, read once per workgroup (all 8 waves use every value)
        __shared__ float2 s_dz[GS_WAVE][GS_PB_G];
        // units j = 4 wave + u: W1 rows, b1, W2 of both outputs as 16-byte loads
        float w[4][3], bb[4], v0[4], v1[4];
        if (valid) {
            const float4* W1v = reinterpret_cast<const float4*>(a.mlp + (size_t)gg * 96 + 12 * wave);
            const float4 wa = W1v[0], wb = W1v[1], wc = W1v[2];
            const float wl[12] = {wa.x, wa.y, wa.z, wa.w, wb.x, wb.y, wb.z, wb.w, wc.x, wc.y, wc.z, wc.w};
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int i = 0; i < 3; ++i) w[u][i] = wl[3 * u + i];
            const float4 bq = *reinterpret_cast<const float4*>(a.mlp + (size_t)P * 96 + (size_t)gg * 32 + 4 * wave);
            const float4 va = *reinterpret_cast<const float4*>(a.mlp + (size_t)P * 128 + (size_t)gg * 64 + 4 * wave);
            const float4 vb =
                *reinterpret_cast<const float4*>(a.mlp + (size_t)P * 128 + (size_t)gg * 64 + 32 + 4 * wave);
            bb[0] = bq.x; bb[1] = bq.y; bb[2] = bq.z; bb[3] = bq.w;
            v0[0] = va.x; v0[1] = va.y; v0[2] = va.z; v0[3] = va.w;
            v1[0] = vb.x; v1[1] = vb.y; v1[2] = vb.z; v1[3] = vb.w;
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) { w[u][0] = w[u][1] = w[u][2] = bb[u] = v0[u] = v1[u] = 0.0f; }
        }
        float gw[4][3], gbias[4], gv0[4], gv1[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) { gw[u][0] = gw[u][1] = gw[u][2] = gbias[u] = gv0[u] = gv1[u] = 0.0f; }
        for (int rb = 0; rb < (int)sc.R; rb += GS_WAVE) {
            const int nrb = min(GS_WAVE, (int)sc.R - rb);
            __syncthreads();   // the previous chunk's receivers are consumed
            if (tid < nrb) {
                const int r = (int)sc.r0 + rb + tid;
                s_roff[tid] = rx_map(d, r).roff;
                s_rx[3 * tid] = a.rx[3 * r];
                s_rx[3 * tid + 1] = a.rx[3 * r + 1];
                s_rx[3 * tid + 2] = a.rx[3 * r + 2];
            }
            __syncthreads();
            // the chunk's dL/d(MLP output) into LDS: wave w loads receivers
            // w, w + 8, ... (64 consecutive Gaussians each: coalesced), all of
            // a thread's loads issued before its stores
            {
                constexpr int NK = GS_WAVE / (GS_PB_THREADS / GS_WAVE);
                const int gc = valid ? gl : 0;
                float2 v[NK];
#pragma unroll
                for (int q = 0; q < NK; ++q) {
                    const int k = wave + q * (GS_PB_THREADS / GS_WAVE);
                    v[q] = k < nrb ? a.dz[(size_t)s_roff[k] + gc] : make_float2(0.0f, 0.0f);
                }
#pragma unroll
                for (int q = 0; q < NK; ++q) {
                    const int k = wave + q * (GS_PB_THREADS / GS_WAVE);
                    s_dz[k][lane] = valid ? v[q] : make_float2(0.0f, 0.0f);
                }
            }
            __syncthreads();
            for (int k = 0; k < nrb; ++k) {
                const float2 dz = s_dz[k][lane];
                const float rx0 = s_rx[3 * k], rx1 = s_rx[3 * k + 1], rx2 = s_rx[3 * k + 2];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    float hs = bb[u];
                    hs = fmaf(w[u][0], rx0, hs);
                    hs = fmaf(w[u][1], rx1, hs);
                    hs = fmaf(w[u][2], rx2, hs);
                    // leaky ReLU (slope 0.01) and its derivative share one select:
                    // hs * 1 and dh * 1 are exact
                    const float slope = hs > 0.0f ? 1.0f : 0.01f;
                    const float h = hs * slope;
                    gv0[u] = fmaf(dz.x, h, gv0[u]);
                    gv1[u] = fmaf(dz.y, h, gv1[u]);
                    const float dh = fmaf(dz.y, v1[u], dz.x * v0[u]);
                    const float dp = dh * slope;
                    gw[u][0] = fmaf(dp, rx0, gw[u][0]);
                    gw[u][1] = fmaf(dp, rx1, gw[u][1]);
                    gw[u][2] = fmaf(dp, rx2, gw[u][2]);
                    gbias[u] += dp;
                }
            }
        }
        if (valid) {
            float4* gW1 = reinterpret_cast<float4*>(a.dmlp + (size_t)gg * 96 + 12 * wave);
            gW1[0] = make_float4(gw[0][0], gw[0][1], gw[0][2], gw[1][0]);
            gW1[1] = make_float4(gw[1][1], gw[1][2], gw[2][0], gw[2][1]);
            gW1[2] = make_float4(gw[2][2], gw[3][0], gw[3][1], gw[3][2]);
            *reinterpret_cast<float4*>(a.dmlp + (size_t)P * 96 + (size_t)gg * 32 + 4 * wave) =
                make_float4(gbias[0], gbias[1], gbias[2], gbias[3]);
            *reinterpret_cast<float4*>(a.dmlp + (size_t)P * 128 + (size_t)gg * 64 + 4 * wave) =
                make_float4(gv0[0], gv0[1], gv0[2], gv0[3]);
            *reinterpret_cast<float4*>(a.dmlp + (size_t)P * 128 + (size_t)gg * 64 + 32 + 4 * wave) =
                make_float4(gv1[0], gv1[1], gv1[2], gv1[3]);
        }
    }
    if (wave == 0 && valid) pre_bwd_finish(d, a, sc, gl, gg);
}

hipError_t launch_preprocess_bwd(const Dims& d, const PreBwdArgs& a, hipStream_t s)
{
    if (d.P == 0 || d.R == 0) return hipSuccess;
    const dim3 g1((d.P + GS_BWD_THREADS - 1) / GS_BWD_THREADS, d.G);
    if (a.mode == GRAD_REFERENCE)
        hipLaunchKernelGGL(k_pre_bwd_geo<GRAD_REFERENCE>, g1, dim3(GS_BWD_THREADS), 0, s, d, a);
    else
        hipLaunchKernelGGL(k_pre_bwd_geo<GRAD_EXACT>, g1, dim3(GS_BWD_THREADS), 0, s, d, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_pre_bwd_param, dim3((d.P + GS_PB_G - 1) / GS_PB_G, d.S), dim3(GS_PB_THREADS), 0, s, d, a);
    return hipGetLastError();
}

}  // namespace gs
