// The training step around the rasterizer (gaussian.py:171-551,
// train.py:364-566): the optimizer step over the raw parameters with their
// activations, the densification statistics, the opacity reset and the
// densify / prune compaction.
//
// Raw parameters and their activations (gaussian.py:33-38):
//   centre [P,3]; log-scale [P,3] -> scale = exp; rotation [P,4] -> unit
//   quaternion q / max(|q|, 1e-12); logit [P] -> opacity = sigmoid; the
//   emission MLPs [P*194] (W1 | b1 | W2 | b2 sections; no activation).
// The rasterizer's backward gives gradients with respect to the activated
// tensors.  The optimizer kernels carry them through the activations, run
// torch.optim.Adam's update and write the raw parameters and, for the next
// forward, their activations.
#include <cmath>

#include "gs_device.h"
#include "gs_kernels.h"

namespace gs {

// ---------------------------------------------------------------- Adam
// Per-step scalars of every group, from the device counters: counters[0] is
// the iteration being completed (the reference's 1-based `iteration`),
// counters[1 + g] the number of steps group g has taken.  Only the groups in
// `groups` step (a parameter the reference has just replaced has no
// gradient and is skipped by torch.optim.Adam, train.py:564 after
// gaussian.py:233-305, 339-393).
__global__ void k_adam_prep(AdamCfg c, int groups, uint32_t* counters, AdamStep* out)
{
    // thread g < AG_N: group g (its FP64 powers and, for the centres, the
    // schedule run in parallel); thread 0 then advances the iteration
    const int g = threadIdx.x;
    const double it = (double)counters[0];
    if (g < AG_N) {
        float step_size = 0.0f, bc2 = 1.0f;
        if ((groups >> g) & 1) {
            const uint32_t t = counters[1 + g] + 1u;
            counters[1 + g] = t;
            double lr = c.lr[g];
            if (g == AG_CENTRE) {
                // get_expon_lr_func (utils.py:46-79) at the iteration number
                if (it < 0.0 || (c.lr_init == 0.0 && c.lr_final == 0.0)) {
                    lr = 0.0;
                } else {
                    double delay = 1.0;
                    if (c.lr_delay_steps > 0)
                        delay = c.lr_delay_mult + (1.0 - c.lr_delay_mult) *
                                                      sin(0.5 * M_PI * fmin(fmax(it / c.lr_delay_steps, 0.0), 1.0));
                    const double tau = fmin(fmax(it / c.lr_max_steps, 0.0), 1.0);
                    lr = delay * exp(log(c.lr_init) * (1.0 - tau) + log(c.lr_final) * tau);
                }
            }
            // torch.optim.Adam (non-capturable): step_size = lr / (1 - beta1^t),
            // denom = sqrt(v) / sqrt(1 - beta2^t) + eps
            step_size = (float)(lr / (1.0 - pow(c.beta1, (double)t)));
            bc2 = (float)sqrt(1.0 - pow(c.beta2, (double)t));
        }
        out->step_size[g] = step_size;
        out->bc2_sqrt[g] = bc2;
    }
    if (g == 0) {
        out->one_minus_b1 = (float)(1.0 - c.beta1);
        out->b2 = (float)c.beta2;
        out->one_minus_b2 = (float)(1.0 - c.beta2);
        out->eps = (float)c.eps;
    }
    __syncthreads();   // every thread has read the iteration number
    if (g == 0) counters[0] = counters[0] + 1u;
}

// one Adam update of x with gradient g and moments m, v (torch's order:
// exp_avg.lerp_(g, 1 - beta1); exp_avg_sq.mul_(beta2).addcmul_(g, g, 1 -
// beta2); x.addcdiv_(exp_avg, sqrt(exp_avg_sq) / bc2_sqrt + eps, -step_size))
__device__ __forceinline__ float adam_update(float x, float g, float& m, float& v, const AdamStep& s, int grp)
{
    m = fmaf(s.one_minus_b1, g - m, m);
    v = s.b2 * v + s.one_minus_b2 * g * g;
    const float denom = sqrtf(v) / s.bc2_sqrt[grp] + s.eps;
    return x + (-s.step_size[grp]) * (m / denom);
}

__device__ __forceinline__ float clamp_keep_nan(float x, float lo, float hi)
{
    return x != x ? x : fminf(fmaxf(x, lo), hi);
}

// Per Gaussian: centre, log-scale, rotation and logit, their activations and
// the centre clamp.  Moments: state [2][11 P] (exp_avg, then exp_avg_sq; each
// centre 3P | log-scale 3P | rotation 4P | logit P).
__global__ void __launch_bounds__(256) k_adam_geom(int P, int groups, AdamCfg c, Gaussians p, GaussianGrads d,
                                                   const AdamStep* stp, uint32_t* counters)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= P) return;
    const AdamStep s = *stp;
    float* m = p.state;
    float* v = p.state + 11 * (size_t)P;
    // centre: NaN gradients are zeroed, infinite ones kept (train.py:441-444);
    // the result is clamped to the scene bounds, NaN kept (train.py:558-562)
    bool nan_centre = false;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const size_t o = 3 * (size_t)i + k;
        float x = p.centre[o];
        if (groups & (1 << AG_CENTRE)) {
            float g = d.centre[o];
            if (g != g) g = 0.0f;
            x = adam_update(x, g, m[o], v[o], s, AG_CENTRE);
        }
        if (c.clamp) x = clamp_keep_nan(x, c.lo[k], c.hi[k]);
        p.centre[o] = x;
        nan_centre |= x != x;
    }
    // the NaN check after the step (train.py:565-566) as a device flag
    if (nan_centre) counters[1 + AG_N] = 1u;
    // log-scale: d/dlog s = d/ds * exp(log s), the forward's exp output
    if (groups & (1 << AG_SCALING)) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const size_t o = 3 * (size_t)i + k;
            const float g = d.scale[o] * p.scale[o];
            const size_t so = 3 * (size_t)P + o;
            const float x = adam_update(p.log_scale[o], g, m[so], v[so], s, AG_SCALING);
            p.log_scale[o] = x;
            p.scale[o] = expf(x);
        }
    }
    // rotation: y = q / n, n = max(|q|, 1e-12); dq = (dy - y (y . dy)) / n,
    // or dy / n where the clamp is active
    if (groups & (1 << AG_ROTATION)) {
        float q[4], dy[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            q[k] = p.rotation[4 * (size_t)i + k];
            dy[k] = d.rotation[4 * (size_t)i + k];
        }
        const float nq = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        const bool clamped = !(nq > 1e-12f);
        const float n = clamped ? 1e-12f : nq;
        float ydy = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k) ydy += (q[k] / n) * dy[k];
        float qn[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float g = clamped ? dy[k] / n : (dy[k] - (q[k] / n) * ydy) / n;
            const size_t so = 6 * (size_t)P + 4 * (size_t)i + k;
            qn[k] = adam_update(q[k], g, m[so], v[so], s, AG_ROTATION);
            p.rotation[4 * (size_t)i + k] = qn[k];
        }
        const float n2 = fmaxf(sqrtf(qn[0] * qn[0] + qn[1] * qn[1] + qn[2] * qn[2] + qn[3] * qn[3]), 1e-12f);
#pragma unroll
        for (int k = 0; k < 4; ++k) p.rotation_n[4 * (size_t)i + k] = qn[k] / n2;
    }
    // logit: d/dlogit = d/do * (1 - o) * o, o the forward's sigmoid output
    if (groups & (1 << AG_OPACITY)) {
        const float o = p.opacity[i];
        const float g = d.opacity[i] * (1.0f - o) * o;
        const size_t so = 10 * (size_t)P + i;
        const float x = adam_update(p.logit[i], g, m[so], v[so], s, AG_OPACITY);
        p.logit[i] = x;
        p.opacity[i] = 1.0f / (1.0f + expf(-x));
    }
}

// A flat array of one group (the emission MLPs; the confidence network).
// Moments: state [2][n].
__global__ void __launch_bounds__(256) k_adam_flat(size_t n, int grp, float* x, const float* g, float* state,
                                                   const AdamStep* stp)
{
    const AdamStep s = *stp;
    float* m = state;
    float* v = state + n;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        x[i] = adam_update(x[i], g[i], m[i], v[i], s, grp);
}

// The same over float4s (16-byte aligned arrays): a quarter of the memory
// instructions, more bytes in flight per wave.
__global__ void __launch_bounds__(256) k_adam_flat4(size_t n4, int grp, float4* x, const float4* g, float4* m,
                                                    float4* v, const AdamStep* stp)
{
    const AdamStep s = *stp;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)gridDim.x * 256) {
        float4 xi = x[i], mi = m[i], vi = v[i];
        const float4 gi = g[i];
        xi.x = adam_update(xi.x, gi.x, mi.x, vi.x, s, grp);
        xi.y = adam_update(xi.y, gi.y, mi.y, vi.y, s, grp);
        xi.z = adam_update(xi.z, gi.z, mi.z, vi.z, s, grp);
        xi.w = adam_update(xi.w, gi.w, mi.w, vi.w, s, grp);
        x[i] = xi;
        m[i] = mi;
        v[i] = vi;
    }
}

// ---------------------------------------------------- densification stats
// add_densification_stats (gaussian.py:538-551) and the max-radius update
// (train.py:510-516) for the R receivers of a batch: for each receiver that
// sees Gaussian i (radius > 0), accum += |dL/dmean2D[:2]| * grad_scale,
// denom += 1, max_radius = max(max_radius, radius).  grad_scale converts the
// gradient of a batch loss to the per-view gradient the reference takes
// (R for a mean over R spectra).
// Workgroup = 64 Gaussians x 4 receiver groups: thread (g, j) sums the
// receivers j, j + 4, ... of Gaussian g (their loads four at a time, all in
// flight together), then the four groups' sums are added in group order, so
// the result does not depend on timing.
__global__ void __launch_bounds__(256) k_densify_stats(int P, int R, float grad_scale, const float* dmeans2d,
                                                       const int* radii, float* accum, float* denom,
                                                       float* max_radius)
{
    __shared__ float s_a[4][64], s_d[4][64], s_m[4][64];
    const int gl = threadIdx.x & 63, grp = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + gl;
    float a = 0.0f, dn = 0.0f, mr = 0.0f;
    if (i < P) {
        for (int r0 = grp; r0 < R; r0 += 16) {
            int rad[4];
            float gx[4], gy[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int r = r0 + 4 * u;
                const size_t o = (size_t)min(r, R - 1) * P + i;
                rad[u] = r < R ? radii[o] : 0;
                gx[u] = dmeans2d[3 * o];
                gy[u] = dmeans2d[3 * o + 1];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (rad[u] > 0) {
                    a += sqrtf(gx[u] * gx[u] + gy[u] * gy[u]) * grad_scale;
                    dn += 1.0f;
                    mr = fmaxf(mr, (float)rad[u]);
                }
            }
        }
    }
    s_a[grp][gl] = a;
    s_d[grp][gl] = dn;
    s_m[grp][gl] = mr;
    __syncthreads();
    if (grp == 0 && i < P) {
        const float sa = ((s_a[0][gl] + s_a[1][gl]) + s_a[2][gl]) + s_a[3][gl];
        const float sd = ((s_d[0][gl] + s_d[1][gl]) + s_d[2][gl]) + s_d[3][gl];
        const float sm = fmaxf(fmaxf(s_m[0][gl], s_m[1][gl]), fmaxf(s_m[2][gl], s_m[3][gl]));
        accum[i] += sa;
        denom[i] += sd;
        max_radius[i] = fmaxf(max_radius[i], sm);
    }
}

// reset_opacity (gaussian.py:223-231): logit = inverse_sigmoid(min(o, 0.01))
// with torch.nan_to_num(nan=0, posinf=1, neginf=0); the logit's Adam moments
// restart at zero (replace_tensor_to_optimizer, gaussian.py:233-253).
__global__ void __launch_bounds__(256) k_opacity_reset(int P, Gaussians p)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= P) return;
    const float o = fminf(p.opacity[i], 0.01f);
    float z = logf(o / (1.0f - o));
    if (z != z) z = 0.0f;
    else if (z == __builtin_huge_valf()) z = 1.0f;
    else if (z == -__builtin_huge_valf()) z = 0.0f;
    p.logit[i] = z;
    p.opacity[i] = 1.0f / (1.0f + expf(-z));
    p.state[10 * (size_t)P + i] = 0.0f;
    p.state[21 * (size_t)P + i] = 0.0f;
}

// ------------------------------------------------- stable list compaction
// Items 0..n-1 each belong to up to NC lists (bit c of f(i)); the lists are
// written one after another, each in item order: out[base_c + rank] = i with
// base_c the total length of the lists before c.  Three passes: per-workgroup
// counts, one workgroup's exclusive scan over them, and the ordered scatter.
constexpr int PART_ITEMS = 1024;   // items per workgroup: 4 rounds of 256

inline unsigned part_blocks(uint32_t n) { return (unsigned)((n + PART_ITEMS - 1) / PART_ITEMS); }

template <int NC, class F>
__global__ void __launch_bounds__(256) k_part_count(uint32_t n, F f, uint32_t* bc)
{
    __shared__ uint32_t s_cnt[4][NC];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t cnt[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) cnt[c] = 0;
    for (int it = 0; it < 4; ++it) {
        const uint32_t i = blockIdx.x * PART_ITEMS + it * 256 + threadIdx.x;
        const uint32_t m = i < n ? f(i) : 0u;
#pragma unroll
        for (int c = 0; c < NC; ++c) cnt[c] += (uint32_t)__popcll(ballot((m >> c) & 1u));
    }
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < NC; ++c) s_cnt[wave][c] = cnt[c];
    __syncthreads();
    if (threadIdx.x < NC) {
        const int c = threadIdx.x;
        bc[(size_t)blockIdx.x * NC + c] = s_cnt[0][c] + s_cnt[1][c] + s_cnt[2][c] + s_cnt[3][c];
    }
}

// bc[b][c] -> exclusive prefix over b; totals[c] = the sum (one workgroup)
template <int NC>
__global__ void __launch_bounds__(1024) k_part_scan(uint32_t nb, uint32_t* bc, uint32_t* totals)
{
    __shared__ uint32_t s_w[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c = 0; c < NC; ++c) {
        uint32_t carry = 0;
        for (uint32_t b0 = 0; b0 < nb; b0 += 1024) {
            const uint32_t b = b0 + threadIdx.x;
            const uint32_t x = b < nb ? bc[(size_t)b * NC + c] : 0u;
            // inclusive scan in the wave
            uint32_t y = x;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t t = __shfl_up(y, off);
                if (lane >= off) y += t;
            }
            if (lane == 63) s_w[wave] = y;
            __syncthreads();
            uint32_t before = 0, all = 0;
            for (int w = 0; w < 16; ++w) {
                const uint32_t t = s_w[w];
                before += w < wave ? t : 0u;
                all += t;
            }
            if (b < nb) bc[(size_t)b * NC + c] = carry + before + y - x;
            carry += all;
            __syncthreads();
        }
        if (threadIdx.x == 0) totals[c] = carry;
    }
}

template <int NC, class F>
__global__ void __launch_bounds__(256) k_part_scatter(uint32_t n, F f, const uint32_t* bc, const uint32_t* totals,
                                                      uint32_t* out)
{
    __shared__ uint32_t s_w[4][NC];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t base[NC];
    uint32_t list0 = 0;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        base[c] = list0 + bc[(size_t)blockIdx.x * NC + c];
        list0 += totals[c];
    }
    for (int it = 0; it < 4; ++it) {
        const uint32_t i = blockIdx.x * PART_ITEMS + it * 256 + threadIdx.x;
        const uint32_t m = i < n ? f(i) : 0u;
        uint64_t bl[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            bl[c] = ballot((m >> c) & 1u);
            if (lane == 0) s_w[wave][c] = (uint32_t)__popcll(bl[c]);
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            uint32_t before = 0, all = 0;
            for (int w = 0; w < 4; ++w) {
                before += w < wave ? s_w[w][c] : 0u;
                all += s_w[w][c];
            }
            if ((m >> c) & 1u) out[base[c] + before + (uint32_t)__popcll(bl[c] & lt)] = i;
            base[c] += all;
        }
        __syncthreads();
    }
}

template <int NC, class F>
hipError_t stable_lists(uint32_t n, F f, uint32_t* bc, uint32_t* totals, uint32_t* out, hipStream_t s)
{
    const unsigned nb = std::max(1u, part_blocks(n));
    hipLaunchKernelGGL((k_part_count<NC, F>), dim3(nb), dim3(256), 0, s, n, f, bc);
    hipLaunchKernelGGL((k_part_scan<NC>), dim3(1), dim3(1024), 0, s, (uint32_t)nb, bc, totals);
    hipLaunchKernelGGL((k_part_scatter<NC, F>), dim3(nb), dim3(256), 0, s, n, f, (const uint32_t*)bc,
                       (const uint32_t*)totals, out);
    return hipGetLastError();
}

// ------------------------------------------------------- densify and prune
// densify_and_prune (gaussian.py:510-535) with densify_and_clone (493-508),
// densify_and_split (450-491), densification_postfix (395-429) and
// prune_points (308-337).  flags[i]: 1 clone (mean gradient >= threshold,
// max scale <= clone_max_scale), 2 split (same gradient test, max scale
// above), 0 neither.
__global__ void __launch_bounds__(256) k_densify_flags(int P, DensifyCfg c, Gaussians p, uint8_t* flags)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= P) return;
    uint8_t f = 0;
    if (c.densify) {
        float g = p.accum[i] / p.denom[i];
        if (g != g) g = 0.0f;                                 // grads[grads.isnan()] = 0
        if (g >= c.grad_threshold) {
            const float smax = fmaxf(fmaxf(p.scale[3 * i], p.scale[3 * i + 1]), p.scale[3 * i + 2]);
            f = smax <= c.clone_max_scale ? 1 : 2;
        }
    }
    flags[i] = f;
}

// list 0: originals that stay (not split); list 1: clones; list 2: split
struct DensifyLists {
    const uint8_t* flags;
    __device__ uint32_t operator()(uint32_t i) const
    {
        const uint8_t f = flags[i];
        return (f != 2 ? 1u : 0u) | (f == 1 ? 2u : 0u) | (f == 2 ? 4u : 0u);
    }
};

struct KeepList {
    const uint8_t* keep;
    __device__ uint32_t operator()(uint32_t i) const { return keep[i]; }
};

// The candidate list after cloning and splitting, in the reference's order:
// the originals that are not split, the clones, then the split children
// (the first child of every split original, then the second), each run in
// original order.  kind: 0 original, 1 clone, 2 child; child = index among
// the children (the row of the noise).
struct Candidate { uint32_t src; int kind; uint32_t child; };

__device__ __forceinline__ Candidate candidate(const DensifyPlan& pl, const uint32_t* order, uint32_t j)
{
    Candidate c;
    if (j < pl.n_keep_orig) {
        c.src = order[j]; c.kind = 0; c.child = 0;
    } else if (j < pl.n_keep_orig + pl.n_clone) {
        c.src = order[j]; c.kind = 1; c.child = 0;
    } else {
        c.child = j - pl.n_keep_orig - pl.n_clone;
        const uint32_t rank = c.child >= pl.n_split ? c.child - pl.n_split : c.child;
        c.src = order[pl.n_keep_orig + pl.n_clone + rank];
        c.kind = 2;
    }
    return c;
}

// Centre, log-scale and scale of a candidate, and the prune test.  A split
// child (N = 2) has centre R(q) (s * z) + mu with z its noise row and R(q)
// the rotation matrix of the normalised quaternion (build_rotation,
// utils.py:127-148), log-scale log(s / (0.6 N)).  Prune (gaussian.py:519-
// 530): opacity < min_opacity, max scale > prune_max_scale, centre outside
// the bounds.  The screen-size test reads max_radii2D after
// densification_postfix has zeroed it (gaussian.py:427-429), so it never
// fires and is not evaluated.
__device__ void candidate_geom(const DensifyCfg& c, const Gaussians& p, const Candidate& cd, const float* noise,
                               float* mu, float* ls, float* sc, bool& keep)
{
    const uint32_t s = cd.src;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        mu[k] = p.centre[3 * (size_t)s + k];
        ls[k] = p.log_scale[3 * (size_t)s + k];
        sc[k] = p.scale[3 * (size_t)s + k];
    }
    if (cd.kind == 2) {
        float q[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) q[k] = p.rotation[4 * (size_t)s + k];
        const float n = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        const float r = q[0] / n, x = q[1] / n, y = q[2] / n, z = q[3] / n;
        const float R[3][3] = {{1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y - r * z), 2.0f * (x * z + r * y)},
                               {2.0f * (x * y + r * z), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z - r * x)},
                               {2.0f * (x * z - r * y), 2.0f * (y * z + r * x), 1.0f - 2.0f * (x * x + y * y)}};
        const float* zz = noise + 3 * (size_t)cd.child;
        const float smp[3] = {sc[0] * zz[0], sc[1] * zz[1], sc[2] * zz[2]};
#pragma unroll
        for (int k = 0; k < 3; ++k) mu[k] = (R[k][0] * smp[0] + R[k][1] * smp[1] + R[k][2] * smp[2]) + mu[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            ls[k] = logf(sc[k] / 1.2f);
            sc[k] = expf(ls[k]);
        }
    }
    const float o = p.opacity[s];
    keep = !(o < c.min_opacity);
    keep = keep && !(fmaxf(fmaxf(sc[0], sc[1]), sc[2]) > c.prune_max_scale);
    if (c.use_bounds)
#pragma unroll
        for (int k = 0; k < 3; ++k) keep = keep && !(mu[k] < c.lo[k]) && !(mu[k] > c.hi[k]);
}

// keep[j] for j < n_max (zero past the candidate count, which is read from
// the totals of the list pass)
__global__ void __launch_bounds__(256) k_densify_keep(uint32_t n_max, DensifyCfg c, Gaussians p, const float* noise,
                                                      const uint32_t* totals, const uint32_t* order, uint8_t* keep)
{
    const uint32_t j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n_max) return;
    DensifyPlan pl;
    pl.P = 0;
    pl.P_new = 0;
    pl.n_keep_orig = totals[0];
    pl.n_clone = totals[1];
    pl.n_split = totals[2];
    pl.n_cand = pl.n_keep_orig + pl.n_clone + 2 * pl.n_split;
    if (j >= pl.n_cand) {
        keep[j] = 0;
        return;
    }
    const Candidate cd = candidate(pl, order, j);
    float mu[3], ls[3], sc[3];
    bool k;
    candidate_geom(c, p, cd, noise, mu, ls, sc, k);
    keep[j] = k ? 1 : 0;
}

// Per output Gaussian o: centre, log-scale, rotation, logit and their
// activations, Adam moments (kept originals carry theirs, new entries start
// at zero: cat_tensors_to_optimizer, gaussian.py:369-378) and the
// statistics.
__global__ void __launch_bounds__(256) k_densify_gather(DensifyPlan pl, DensifyCfg c, Gaussians p, Gaussians q,
                                                        const float* noise, const uint32_t* order,
                                                        const uint32_t* src_of)
{
    const uint32_t o = blockIdx.x * 256 + threadIdx.x;
    if (o >= pl.P_new) return;
    const Candidate cd = candidate(pl, order, src_of[o]);
    float mu[3], ls[3], sc[3];
    bool keep;
    candidate_geom(c, p, cd, noise, mu, ls, sc, keep);
    const uint32_t s = cd.src;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        q.centre[3 * (size_t)o + k] = mu[k];
        q.log_scale[3 * (size_t)o + k] = ls[k];
        q.scale[3 * (size_t)o + k] = sc[k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        q.rotation[4 * (size_t)o + k] = p.rotation[4 * (size_t)s + k];
        q.rotation_n[4 * (size_t)o + k] = p.rotation_n[4 * (size_t)s + k];
    }
    q.logit[o] = p.logit[s];
    q.opacity[o] = p.opacity[s];
    const size_t Po = pl.P, Pn = pl.P_new;
    const int seg_w[4] = {3, 3, 4, 1};
    size_t s_off = 0, d_off = 0;
#pragma unroll
    for (int sg = 0; sg < 4; ++sg) {
        const int w = seg_w[sg];
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const size_t sb = (size_t)half * 11 * Po + s_off, db = (size_t)half * 11 * Pn + d_off;
            for (int k = 0; k < w; ++k)
                q.state[db + (size_t)w * o + k] = cd.kind == 0 ? p.state[sb + (size_t)w * s + k] : 0.0f;
        }
        s_off += (size_t)w * Po;
        d_off += (size_t)w * Pn;
    }
    // densification_postfix zeroes the statistics of every Gaussian before
    // the prune; pruning alone carries them (prune_points, gaussian.py:334-337)
    q.accum[o] = c.densify ? 0.0f : p.accum[s];
    q.denom[o] = c.densify ? 0.0f : p.denom[s];
    q.max_radius[o] = c.densify ? 0.0f : p.max_radius[s];
}

// The emission MLPs and their moments, one thread per float: element k of
// Gaussian o's 194 (sections W1 96 | b1 32 | W2 64 | b2 2, each [P, width]).
__global__ void __launch_bounds__(256) k_densify_gather_mlp(DensifyPlan pl, Gaussians p, Gaussians q,
                                                            const uint32_t* order, const uint32_t* src_of)
{
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (size_t)pl.P_new * 194) return;
    const uint32_t o = (uint32_t)(e / 194);
    const int k = (int)(e - (size_t)o * 194);
    const Candidate cd = candidate(pl, order, src_of[o]);
    int k0, w;
    if (k < 96) { k0 = 0; w = 96; }
    else if (k < 128) { k0 = 96; w = 32; }
    else if (k < 192) { k0 = 128; w = 64; }
    else { k0 = 192; w = 2; }
    const size_t Po = pl.P, Pn = pl.P_new;
    const size_t si = (size_t)k0 * Po + (size_t)w * cd.src + (k - k0);
    const size_t di = (size_t)k0 * Pn + (size_t)w * o + (k - k0);
    q.mlp[di] = p.mlp[si];
    q.mlp_state[di] = cd.kind == 0 ? p.mlp_state[si] : 0.0f;
    q.mlp_state[194 * Pn + di] = cd.kind == 0 ? p.mlp_state[194 * Po + si] : 0.0f;
}

// The activations from the raw parameters (gaussian.py:86-96).
__global__ void __launch_bounds__(256) k_activate(int P, Gaussians p)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= P) return;
#pragma unroll
    for (int k = 0; k < 3; ++k) p.scale[3 * (size_t)i + k] = expf(p.log_scale[3 * (size_t)i + k]);
    float q[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = p.rotation[4 * (size_t)i + k];
    const float n = fmaxf(sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]), 1e-12f);
#pragma unroll
    for (int k = 0; k < 4; ++k) p.rotation_n[4 * (size_t)i + k] = q[k] / n;
    p.opacity[i] = 1.0f / (1.0f + expf(-p.logit[i]));
}

// ---------------------------------------------------------------- launchers
static inline unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_activate(int P, const Gaussians& p, hipStream_t s)
{
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(k_activate, dim3(blocks_for(P)), dim3(256), 0, s, P, p);
    return hipGetLastError();
}

hipError_t launch_adam_step(int P, int groups, const AdamCfg& c, const Gaussians& p, const GaussianGrads& d,
                            size_t n_conf, float* conf, const float* dconf, float* conf_state, uint32_t* counters,
                            AdamStep* step, hipStream_t s)
{
    hipLaunchKernelGGL(k_adam_prep, dim3(1), dim3(64), 0, s, c, groups, counters, step);
    if (P > 0)
        hipLaunchKernelGGL(k_adam_geom, dim3(blocks_for(P)), dim3(256), 0, s, P, groups, c, p, d,
                           (const AdamStep*)step, counters);
    const size_t n = (size_t)P * 194;
    if (n > 0 && (groups & (1 << AG_EMISSION))) {
        const bool vec = n % 4 == 0 && ((uintptr_t)p.mlp & 15) == 0 && ((uintptr_t)d.mlp & 15) == 0 &&
                         ((uintptr_t)p.mlp_state & 15) == 0;
        if (vec)
            hipLaunchKernelGGL(k_adam_flat4, dim3((unsigned)std::min<size_t>(blocks_for(n / 4), 8192)), dim3(256), 0,
                               s, n / 4, (int)AG_EMISSION, (float4*)p.mlp, (const float4*)d.mlp,
                               (float4*)p.mlp_state, (float4*)(p.mlp_state + n), (const AdamStep*)step);
        else
            hipLaunchKernelGGL(k_adam_flat, dim3((unsigned)std::min<size_t>(blocks_for(n), 8192)), dim3(256), 0, s,
                               n, (int)AG_EMISSION, p.mlp, d.mlp, p.mlp_state, (const AdamStep*)step);
    }
    if (n_conf > 0 && (groups & (1 << AG_CONFIDENCE)))
        hipLaunchKernelGGL(k_adam_flat, dim3(blocks_for(n_conf)), dim3(256), 0, s, n_conf, (int)AG_CONFIDENCE, conf,
                           dconf, conf_state, (const AdamStep*)step);
    return hipGetLastError();
}

hipError_t launch_densify_stats(int P, int R, float grad_scale, const float* dmeans2d, const int* radii,
                                const Gaussians& p, hipStream_t s)
{
    if (P == 0 || R == 0) return hipSuccess;
    hipLaunchKernelGGL(k_densify_stats, dim3((unsigned)((P + 63) / 64)), dim3(256), 0, s, P, R, grad_scale, dmeans2d, radii,
                       p.accum, p.denom, p.max_radius);
    return hipGetLastError();
}

hipError_t launch_opacity_reset(int P, const Gaussians& p, hipStream_t s)
{
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(k_opacity_reset, dim3(blocks_for(P)), dim3(256), 0, s, P, p);
    return hipGetLastError();
}

hipError_t launch_densify_lists(int P, const DensifyCfg& c, const Gaussians& p, const float* noise,
                                const DensifyWork& w, hipStream_t s)
{
    hipLaunchKernelGGL(k_densify_flags, dim3(blocks_for(P)), dim3(256), 0, s, P, c, p, w.flags);
    hipError_t e = stable_lists<3>((uint32_t)P, DensifyLists{w.flags}, w.block_counts, w.totals, w.order, s);
    if (e != hipSuccess) return e;
    const uint32_t n_max = 2u * (uint32_t)P;
    hipLaunchKernelGGL(k_densify_keep, dim3(blocks_for(n_max)), dim3(256), 0, s, n_max, c, p, noise,
                       (const uint32_t*)w.totals, (const uint32_t*)w.order, w.keep);
    return stable_lists<1>(n_max, KeepList{w.keep}, w.block_counts, w.totals + 3, w.src_of, s);
}

hipError_t launch_densify_gather(const DensifyPlan& pl, const DensifyCfg& c, const Gaussians& p, const Gaussians& q,
                                 const float* noise, const DensifyWork& w, hipStream_t s)
{
    if (pl.P_new == 0) return hipSuccess;
    hipLaunchKernelGGL(k_densify_gather, dim3(blocks_for(pl.P_new)), dim3(256), 0, s, pl, c, p, q, noise,
                       (const uint32_t*)w.order, (const uint32_t*)w.src_of);
    hipLaunchKernelGGL(k_densify_gather_mlp, dim3(blocks_for((size_t)pl.P_new * 194)), dim3(256), 0, s, pl, p, q,
                       (const uint32_t*)w.order, (const uint32_t*)w.src_of);
    return hipGetLastError();
}

}  // namespace gs
