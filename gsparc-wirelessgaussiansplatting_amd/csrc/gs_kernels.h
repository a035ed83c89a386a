// gs_kernels.h -- internal launchers of the gfx950 kernels (C++ linkage; the
// exported C ABI lives in gsparc_abi.cpp and include/gsparc.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gs_csi.h"

namespace gs {

// A batch of S scenes (independent Gaussian sets) with R receivers in all.
// Scene s owns Gaussians [g0_s, g0_s + P_s) of the packed parameter arrays
// and R_s consecutive receivers; every (receiver, Gaussian-of-its-scene)
// pair owns one record (splat, emission, depth, radius, sort keys,
// gradients), receiver r's records starting at roff_r.  Receivers are
// processed by the preprocess kernels in groups of up to GS_RC of one scene.
//
// tab == nullptr is the uniform batch (one scene, every receiver sees all P
// Gaussians: roff_r = r*P); otherwise tab is the device copy of the scene
// table (gsparc_scene_table_build):
//   [0..7]                 S, R, G, Pmax, Ptot, Nrec, 0, 0
//   receiver r, 4 words    roff, g0, P, scene          at GS_TAB_HDR + 4r
//   group y, 4 words       r0, nr, scene, part offset  after the receivers
//   scene s, 8 words       g0, P, r0, R, group0, ngroups, 0, 0  after the groups
#define GS_RC 16           // receivers per preprocess group
#define GS_TAB_HDR 8

struct Dims {
    int P, R, W, H;   // P: Gaussians per receiver (the largest scene's count)
    int gx, gy, T;    // tile grid
    int NB;           // blocks of 256 Gaussians
    int Ptot;         // Gaussians in the packed parameter arrays
    int G;            // preprocess receiver groups
    int S;            // scenes
    size_t Nrec;      // (receiver, Gaussian) records
    size_t Npart;     // preprocess-backward partial-sum slots (host side)
    const uint32_t* tab;
};

// depth sort: 256-thread blocks of 8 elements per thread
#define GS_RS_THREADS 256
#define GS_RS_ITEMS 8
#define GS_RS_TILE (GS_RS_THREADS * GS_RS_ITEMS)
// binning chunks: at least this many sorted Gaussians per chunk; per-tile
// counters of a chunk live in LDS up to this many tiles (90x360 has 138,
// 1080x1920 has 8160), in the global count matrix beyond
#define GS_CHUNK_MIN 256
#define GS_CHUNK_LDS_TILES 8192
#define GS_SCATTER_WAVES 8        // independent waves (chunks) per ordered-scatter workgroup
#define GS_ROW_PLACE_MAX_T 1024   // counts and placement per (chunk, tile row) up to this many tiles
#define GS_RP_WAVES 4              // waves per k_row_place workgroup
#define GS_SEGSCAN_LDS 15360      // (receiver, tile) counts scanned through LDS by k_segscan_lds

inline Dims make_dims(int P, int R, int W, int H)
{
    Dims d;
    d.P = P; d.R = R; d.W = W; d.H = H;
    d.gx = (W + 15) / 16; d.gy = (H + 15) / 16; d.T = d.gx * d.gy;
    d.NB = (P + 255) / 256;
    d.Ptot = P;
    d.G = (R + GS_RC - 1) / GS_RC;
    d.S = 1;
    d.Nrec = (size_t)R * P;
    d.Npart = (size_t)d.G * P;
    d.tab = nullptr;
    return d;
}

struct RxMap {
    uint32_t roff, g0, P, scene;
};
struct GroupMap {
    int r0, nr, scene;
    uint32_t part;    // offset of the group's per-Gaussian partial sums
};
struct SceneMap {
    uint32_t g0, P, r0, R, group0, ngroups;
};

__device__ __forceinline__ RxMap rx_map(const Dims& d, int r)
{
    if (!d.tab) return {(uint32_t)r * (uint32_t)d.P, 0u, (uint32_t)d.P, 0u};
    const uint32_t* e = d.tab + GS_TAB_HDR + 4 * r;
    return {e[0], e[1], e[2], e[3]};
}

__device__ __forceinline__ GroupMap group_map(const Dims& d, int y)
{
    if (!d.tab) return {y * GS_RC, min(GS_RC, d.R - y * GS_RC), 0, (uint32_t)y * (uint32_t)d.P};
    const uint32_t* e = d.tab + GS_TAB_HDR + 4 * d.R + 4 * y;
    return {(int)e[0], (int)e[1], (int)e[2], e[3]};
}

__device__ __forceinline__ SceneMap scene_map(const Dims& d, int s)
{
    if (!d.tab) return {0u, (uint32_t)d.P, 0u, (uint32_t)d.R, 0u, (uint32_t)d.G};
    const uint32_t* e = d.tab + GS_TAB_HDR + 4 * d.R + 4 * d.G + 8 * s;
    return {e[0], e[1], e[2], e[3], e[4], e[5]};
}

// scene of packed Gaussian gg (binary search over the scene starts)
__device__ __forceinline__ int scene_of(const Dims& d, int gg)
{
    if (!d.tab) return 0;
    int lo = 0, hi = d.S - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((int)scene_map(d, mid).g0 <= gg) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

struct PreFwdArgs {
    const float* means;      // [Ptot,3]
    const float* scales;     // [Ptot,3]
    const float* rots;       // [Ptot,4]
    const float* opac;       // [Ptot]
    const float* mlp;        // [Ptot*194] or nullptr (emission supplied elsewhere)
    const float* views;      // [R,16] column-major
    const float* tx;         // [S,3]
    const float* rx;         // [R,3]
    float mod;
    float4* splat;           // [Nrec*2]
    float2* sig;             // [Nrec] (written only when mlp != nullptr)
    float* depth;            // [Nrec]
    int* radii;              // [Nrec]
    float* cov3;             // [P*6]
    float* plen;             // [Nrec] transmitter -> Gaussian -> receiver path length, or nullptr
    uint32_t* rect;          // [Nrec] packed tile rect (rect_packed) or nullptr
};
// Tile rects packed in 32 bits: x0 | y0 << 8 | x1 << 16 | y1 << 24 (tile grids
// up to 255 x 255).
__host__ __device__ inline bool rect_packed(const Dims& d) { return d.gx <= 255 && d.gy <= 255; }
hipError_t launch_preprocess_fwd(const Dims& d, const PreFwdArgs& a, hipStream_t s);

struct BinArgs {
    const float4* splat;     // for px, py
    const int* radii;
    const float* depth;
    uint32_t* segcnt;        // [R*T] instances per tile
    uint32_t* segbase;       // [R*T] exclusive prefix of segcnt
    uint2* ranges;           // [R*T]
    uint32_t* total;         // [1]: K
    uint32_t* nvis;          // [R] visible Gaussians per receiver
    uint32_t* skey[2];       // [Nrec] depth-sort ping-pong: depth bits
    uint32_t* sval[2];       // [Nrec]                       Gaussian index
    uint32_t* shist;         // [os_words] depth-sort state: look-back status words, digit counts, tickets
                             //   (or [R][257][rs_blocks(P)] per-tile digit histograms past GS_OS_MAX_TILES)
    uint32_t* lbstat;        // (set by the sort) [R][rs_blocks(P)][256] status words of a pass
    uint32_t* ghist;         // (set by the sort) [R][GS_OS_GHIST] digit counts of the 4 passes, visible count
    uint32_t* tick;          // (set by the sort) [4][R] tile tickets
    uint32_t* ccnt;          // [R*T*nchunks] per-chunk tile counts / bases
    int nchunks, chunk_len;
    uint32_t* plist;         // [Kcap] Gaussian indices in tile-list order
    uint32_t* keys_sorted;   // optional [Kcap] their depth bits
    uint64_t Kcap;
    const uint32_t* rect;    // [Nrec] packed tile rect per record (nullptr: rects from radii and splat)
    uint32_t* srect;         // [Nrec] the same in depth order (written by the last sort pass)
};
int rs_blocks(int P);
#define GS_OS_GHIST 1025
#define GS_OS_MAX_TILES 32   // one-sweep depth sort up to this many 2048-key tiles per receiver
size_t os_words(const Dims& d);   // 32-bit words of BinArgs::shist
int chunk_len_for(const Dims& d);
int chunk_count_for(const Dims& d);
hipError_t launch_depth_sort(const Dims& d, const BinArgs& a, hipStream_t s);
hipError_t launch_chunk_count(const Dims& d, const BinArgs& a, hipStream_t s);
hipError_t launch_chunk_scan(const Dims& d, const BinArgs& a, hipStream_t s);
hipError_t launch_segscan(const Dims& d, const BinArgs& a, hipStream_t s);
// launch_segscan followed by launch_tile_order(mode) on the counts, in one
// launch where the counts fit the LDS scan and the mode has at most 8 lists
hipError_t launch_segscan_order(const Dims& d, const BinArgs& a, int mode, uint32_t* order, hipStream_t s);
hipError_t launch_chunk_scatter(const Dims& d, const BinArgs& a, hipStream_t s);
// Placement of the instances after launch_chunk_count and launch_chunk_scan:
// per (chunk, tile row) waves when row_place_ok(d) (packed rects and at most
// GS_ROW_PLACE_MAX_T tiles), else launch_chunk_scatter.
bool row_place_ok(const Dims& d);
hipError_t launch_bin_place(const Dims& d, const BinArgs& a, hipStream_t s);
// Work order of the blend kernels over the (receiver, tile) pairs
// (item = receiver * T + tile; 0xffffffff = none), longest lists first:
//   ORDER_GLOBAL        one list; consecutive items go to consecutive XCDs
//   ORDER_XCD_MIXED     XCD x takes the receivers r = x (mod 8), all their
//                       tiles longest first (its L2 serves 1/8 of the data)
//   ORDER_XCD_RECEIVER  as ORDER_XCD_MIXED, one receiver after the other
//                       (the L2 holds about one receiver's records at a time)
// The XCD orders are 8 lists of n8 = ceil(R/8)*T slots.  The mode used is
// stored in total[1] for the blend kernels (forward and backward).
enum TileOrder { ORDER_GLOBAL = 1, ORDER_XCD_MIXED = 2, ORDER_XCD_RECEIVER = 3 };
size_t order_length(const Dims& d);
hipError_t launch_tile_order(const Dims& d, int mode, const uint2* ranges, uint32_t* order, uint32_t* total,
                             hipStream_t s);

struct RenderFwdArgs {
    const uint2* ranges;     // [R*T]
    const uint32_t* order;   // work order (see TileOrder; its mode is total[1])
    const uint32_t* plist;
    const float4* splat;     // [Nrec*2]
    const float2* sig;       // [Nrec]
    float* out_re;           // [R*N]
    float* out_im;           // [R*N]
    float* finalT;           // [R*N]
    uint32_t* ncontrib;      // [R*N]
    const uint32_t* total;   // K as computed on the device
    uint64_t Kcap;           // capacity of plist (nothing is read when K > Kcap)
    unsigned long long* counts;  // optional profiling counters (see below)
};
// Profiling counters (u64 x 4, optional), in the reference's terms
// (forward.cu:198-236, backward.cu:547-563), whatever this build culls:
// [0] (pixel, list entry) pairs the reference's forward visits (each pixel's
// list up to its terminating entry, else all of it), [1] pairs its backward
// walks (the contributor counts), [2] / [3] contributing pairs (alpha >=
// 1/255, before termination) as the forward / backward found them.
// pixels per lane of the blend kernels: 1 (8x8 quarter-tile per wave), 2
// (16x8 half tile) or 4 (16x16 tile per wave); requested = 0 picks by the
// number of tiles
int choose_ppl(const Dims& d, int requested);
// list entries staged in LDS per batch of the blend kernels (batch depth):
// 32, 64 or 128; any other request gives 64
int choose_depth(int requested);
hipError_t launch_render_fwd(const Dims& d, const RenderFwdArgs& a, int ppl, int depth, hipStream_t s);

struct RenderBwdArgs {
    const uint2* ranges;
    const uint32_t* order;
    const uint32_t* plist;
    const float4* splat;
    const float2* sig;
    const float* finalT;
    const uint32_t* ncontrib;
    const float* dout_re;    // [R*N]
    const float* dout_im;
    // [Nrec*8] accumulated raw sums over the record's pixels, with
    // D = o G dL/dalpha: D dx, D dy, D dx^2, D dx dy, D dy^2, G dL/dalpha,
    // w dL/dre, w dL/dim (the preprocess backward turns the first five into
    // dL/dmean2D and dL/dconic with the record's conic)
    float* g8;
    int mode;
    const uint32_t* total;
    uint64_t Kcap;
    unsigned long long* counts;  // optional profiling counters
    // deterministic form (inst8 == nullptr: the atomic one): [Kcap*8]
    // per-instance sums, zero-filled by the caller; det_off [Nrec] the first
    // slot of each record's run (launch_det_offsets), radii [Nrec] its radius;
    // the kernel walks whole tiles and the record sums come from
    // launch_det_gather
    float* inst8;
    const uint32_t* det_off;
    const int* radii;
};
hipError_t launch_render_bwd(const Dims& d, const RenderBwdArgs& a, int ppl, int depth, hipStream_t s);

// Deterministic gradient sums: per-record runs of instance slots.
struct DetArgs {
    const float4* splat;     // [Nrec*2]
    const int* radii;        // [Nrec]
    uint32_t* off;           // [Nrec] first slot of each record's run
    uint32_t* bsum;          // [ceil(Nrec/1024)] block sums, then block bases
    const float* inst8;      // [Kcap*8]
    float* g8;               // [Nrec*8] (visible records written, the rest untouched)
    const uint32_t* total;
    uint64_t Kcap;
};
size_t det_scratch_words(const Dims& d);   // off + bsum
hipError_t launch_det_offsets(const Dims& d, const DetArgs& a, hipStream_t s);
hipError_t launch_det_gather(const Dims& d, const DetArgs& a, hipStream_t s);

struct PreBwdArgs {
    const float* means;
    const float* scales;
    const float* rots;
    const float* mlp;        // or nullptr
    const float* views;
    const float* tx;
    const float* rx;
    float mod;
    const int* radii;        // [Nrec]
    const float* g8;         // [Nrec*8]
    const float4* splat;     // [Nrec*2] splat records of the forward (their conics)
    const float2* sig;       // [Nrec] MLP outputs of the forward (used when mlp != nullptr)
    const float* cov3;       // [P*6] cov3D of the forward
    float* dmeans;           // [P,3]
    float* dscales;          // [P,3]
    float* drots;            // [P,4]
    float* dopac;            // [P]
    float* dmlp;             // [P*194] or nullptr
    float* dcov3;            // [P,6] or nullptr
    float* dmeans2d;         // [R*P*3] or nullptr (reference screen-space layout)
    float* dsig;             // [Nrec*2] or nullptr (emission gradient)
    const float* gL;         // [Nrec] dL/d(path length) of the CSI blend, or nullptr
    int mode;
    float2* dz;              // workspace [Nrec]: dL/d(MLP output) per receiver
    float* part;             // workspace [ceil(R/8)*P*12]: per-chunk partial sums
};
hipError_t launch_preprocess_bwd(const Dims& d, const PreBwdArgs& a, hipStream_t s);

// Multi-frequency (CSI) blend: GS_CSI_NC complex channels per pixel, stored
// interleaved [R*N][NC][2]; emission per record s * exp(-j 2 pi f_k L / c)
// (gs_csi.h).  Forward writes out, finalT, ncontrib; backward (exact adjoint)
// accumulates g8 as the single-channel backward does, with dsig the gradient
// of the complex amplitude, plus gL = dL/dL per record.
struct RenderCsiArgs {
    const uint2* ranges;
    const uint32_t* order;
    const uint32_t* plist;
    const float4* splat;
    const float2* sig;       // [Nrec] complex amplitude
    const float* plen;       // [Nrec] path length
    float* out;              // forward: [R*N*NC*2]
    float* finalT;           // forward writes, backward reads
    uint32_t* ncontrib;
    const float* dout;       // backward: [R*N*NC*2]
    float* g8;               // backward: [Nrec*8]
    float* gL;               // backward: [Nrec]
    const uint32_t* total;
    uint64_t Kcap;
    GsCsi csi;
};
hipError_t launch_render_fwd_csi(const Dims& d, const RenderCsiArgs& a, int ppl, hipStream_t s);
hipError_t launch_render_bwd_csi(const Dims& d, const RenderCsiArgs& a, hipStream_t s);

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);

struct SsimArgs {
    int planes, H, W;
    float C1, C2;
    const float* img1;
    const float* img2;
    float* map;              // nullptr: the map is not stored
    float* dmu1;             // nullptr when not training
    float* ds1;
    float* ds12;
    float* part;             // nullptr, or per workgroup (sum of the map, sum of |img1 - img2|)
    int nsum = 2;            // sums per workgroup in part: 2, or 5 (also sum (img1 - img2)^2, img1^2, img2^2)
};
hipError_t launch_ssim_fwd(const SsimArgs& a, hipStream_t s);
int ssim_blocks(int planes, int H, int W);

struct SsimBwdArgs {
    int planes, H, W;
    const float* img1;
    const float* img2;
    const float* dmap;       // nullptr: every pixel's dL/dmap is dmap_const
    const float* dmu1;
    const float* ds1;
    const float* ds12;
    float* dimg1;
    float dmap_const;
    float l1w;               // adds l1w * sign(img1 - img2) to dimg1 (0: nothing)
    const float* plane_scale;  // nullptr, or [planes]: dmap_const and l1w are scaled by plane_scale[plane]
};
hipError_t launch_ssim_bwd(const SsimBwdArgs& a, hipStream_t s);

// The confidence-weighted loss over R spectra (train.py:418-436) from the
// per-workgroup partials of k_ssim_fwd (nbp workgroups per plane):
//   L_r = l1w mean|img_r - tgt_r| + sw (1 - mean SSIM_r),
//   C_r = 1 + exp(w2 . relu(W1 rx_r + b1) + b2) (or 1 without a network),
//   loss = mean_r (C_r L_r - alpha log C_r);
// plane_scale[r] = C_r scales the SSIM backward of plane r; dmlp (may be
// null) = dloss / d(W1 | b1 | w2 | b2).  scratch: 4 floats per plane.
struct ConfLossArgs {
    const float* part;
    int planes, nbp, hidden;
    double inv_np;           // 1 / (H W)
    float l1w, sw, alpha;
    const float* mlp;        // [hidden*3 | hidden | hidden | 1] or nullptr
    const float* rx;         // [planes, 3]
    float* loss;             // [1]
    float* plane_scale;      // [planes]
    float* rec;              // [planes] or nullptr
    float* conf;             // [planes] or nullptr
    float* dmlp;             // [5 hidden + 1] or nullptr
    float* scratch;          // [planes * 4]
};
hipError_t launch_conf_loss(const ConfLossArgs& a, hipStream_t s);

// Evaluation metrics per plane from k_ssim_fwd's five-sum partials (nbp
// workgroups per plane): metrics [planes][EVAL_METRICS] = loss, PSNR, SSIM,
// MAE, MSE, RSSI of img, RSSI(img) - RSSI(target), confidence.
constexpr int EVAL_METRICS = 8;
struct EvalArgs {
    const float* part;
    int planes, nbp, hidden;
    double inv_np;
    float lambda_ssim;
    const float* mlp;        // packed confidence network or nullptr (confidence 1)
    const float* rx;         // [planes, 3]
    float* metrics;
};
hipError_t launch_eval_metrics(const EvalArgs& a, hipStream_t s);

// Shared emission MLP on MFMA (k_mlp_mfma.hip): 68 -> 128 -> 128 -> 1 for every
// (receiver, Gaussian) pair; emission rows r*P + g.
struct Mlp1Args {
    int P, R;
    const float* means;      // [P,3]
    const float* rx;         // [R,3]
    const float* tx;         // [3]
    const float* w1;         // [128,68] (out, in): PE inputs 0..62, receiver features 63..67
    const float* b1;         // [128]
    const float* w2;         // [128,128] (out, in)
    const float* b2;         // [128]
    const float* w3;         // [128]
    const float* b3;         // [1]
    float* pe1;              // [P,128] PE(mu) W1pe^T + b1 (work; written by the forward)
    float* out;              // [R*P,2] emission (a, 0)
};
size_t mlp1_work_floats(int P, int R);
size_t mlp1_grad_floats();
hipError_t launch_mlp1_fwd(const Mlp1Args& a, hipStream_t s);
// grads: dW1 [128*68] | db1 [128] | dW2 [128*128] | db2 [128] | dw3 [128] | db3 [1];
// dmu [P,3] written (the MLP's gradient of the means); work as the forward left it
hipError_t launch_mlp1_bwd(const Mlp1Args& a, const float* dem, float* dmu, float* work, float* grads, hipStream_t s,
                           bool accumulate);

// CSI loss pieces (k_loss_csi.hip): magnitude planes + L1 partials, the
// gradient back to the interleaved map, and the two-part loss reduction.
int csi_loss_blocks(size_t npix);
hipError_t launch_csi_mag(size_t npix, int N, const float* out, const float* tgt, float* magA, float* magB,
                          float* part, hipStream_t s);
hipError_t launch_csi_grad(size_t npix, int N, const float* out, const float* tgt, const float* magA,
                           const float* dmag, float l1w, float* dout, hipStream_t s);
size_t loss_reduce2_scratch();
hipError_t launch_loss_reduce2(const float* ps, int nbs, double inv_s, const float* pl, int nbl, double inv_l,
                               float l1w, float sw, double* stage, float* loss, hipStream_t s);

// ---- the training step around the rasterizer (k_train.hip) ----
// Adam parameter groups, in the bit order of GSPARC_GROUP_*.
enum { AG_CENTRE = 0, AG_SCALING, AG_ROTATION, AG_OPACITY, AG_EMISSION, AG_CONFIDENCE, AG_N };

// The gsparc_gaussians arrays (P entries each).
struct Gaussians {
    float *centre, *log_scale, *rotation, *logit, *mlp;
    float *scale, *rotation_n, *opacity;
    float *state, *mlp_state;
    float *accum, *denom, *max_radius;
};
// Gradients with respect to the activations.
struct GaussianGrads { const float *centre, *scale, *rotation, *opacity, *mlp; };

struct AdamCfg {
    double beta1, beta2, eps;
    double lr[AG_N];                    // per group; lr[AG_CENTRE] comes from the schedule
    double lr_init, lr_final, lr_delay_mult;
    int lr_delay_steps, lr_max_steps;
    int clamp;
    float lo[3], hi[3];
};
// Per-step scalars (k_adam_prep), in the training state after the counters.
struct AdamStep {
    float step_size[AG_N], bc2_sqrt[AG_N];
    float one_minus_b1, b2, one_minus_b2, eps;
};
constexpr size_t TRAIN_STEP_OFFSET = 64;   // byte offset of AdamStep in the training state

struct DensifyCfg {
    int densify;
    float grad_threshold, clone_max_scale, min_opacity, prune_max_scale;
    int use_bounds;
    float lo[3], hi[3];
};
struct DensifyPlan { uint32_t P, n_keep_orig, n_clone, n_split, n_cand, P_new; };
// Work arrays of gsparc_densify_and_prune for P Gaussians: flags [P], keep
// [2P], order [2P], src_of [2P], block_counts [3 * ceil(2P / 1024)],
// totals [4] (the three list lengths, then the kept count).
struct DensifyWork { uint8_t *flags, *keep; uint32_t *order, *src_of, *block_counts, *totals; };

hipError_t launch_activate(int P, const Gaussians& p, hipStream_t s);
hipError_t launch_adam_step(int P, int groups, const AdamCfg& c, const Gaussians& p, const GaussianGrads& d,
                            size_t n_conf, float* conf, const float* dconf, float* conf_state, uint32_t* counters,
                            AdamStep* step, hipStream_t s);
hipError_t launch_densify_stats(int P, int R, float grad_scale, const float* dmeans2d, const int* radii,
                                const Gaussians& p, hipStream_t s);
hipError_t launch_opacity_reset(int P, const Gaussians& p, hipStream_t s);
hipError_t launch_densify_lists(int P, const DensifyCfg& c, const Gaussians& p, const float* noise,
                                const DensifyWork& w, hipStream_t s);
hipError_t launch_densify_gather(const DensifyPlan& pl, const DensifyCfg& c, const Gaussians& p, const Gaussians& q,
                                 const float* noise, const DensifyWork& w, hipStream_t s);

// The multi-frequency kernels' phase factors (cos 2 pi x, sin 2 pi x) of
// cycle fractions x, for their accuracy check (k_render_csi.hip).
hipError_t launch_phase_factors(size_t n, const float* x, float* c, float* s, hipStream_t st);
hipError_t launch_csi_phasors(size_t n, const float* plen, const GsCsi& t, float* out, hipStream_t st);

// ---- the spectrum dataset on the device (k_data.hip) ----
// q: gateway quaternion (x, y, z, w); out [N,16] column-major view matrices.
hipError_t launch_view_matrices(int N, const double* q, const float* rx, float* out, hipStream_t s);
hipError_t launch_gather_batch(int R, int N, int npix, const int* order, int L, const uint32_t* train_state,
                               const float* spectra, const float* rx, const float* views, float* out_spec,
                               float* out_rx, float* out_views, hipStream_t s);

// loss = l1w * sum|img1 - img2| * inv_n + sw * (1 - sum(map) * inv_n) from
// the per-workgroup partials of k_ssim_fwd, in a fixed order.
hipError_t launch_loss_reduce(const float* part, int nb, double inv_n, float l1w, float sw, float* loss,
                              hipStream_t s);

}  // namespace gs
