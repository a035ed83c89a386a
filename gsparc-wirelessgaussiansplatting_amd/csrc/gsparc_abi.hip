// gsparc_abi.hip -- the exported C ABI (include/gsparc.h): buffer layout,
// argument checks and the launch sequence of the forward and backward.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/gsparc.h"
#include "gs_device.h"
#include "gs_kernels.h"

namespace {

thread_local std::string g_err;

// tuning options (gsparc_set_option): pixels per lane of the forward and
// backward blend and the blend work order; 0 = automatic (what the
// autotuner chose for this batch shape, else the built-in rule)
int g_opt_fwd_ppl = 0;
int g_opt_bwd_ppl = 0;
int g_opt_order = 0;
// list entries staged per LDS batch of the forward and backward blend (32,
// 64, 128); 0 = automatic
int g_opt_fwd_depth = 0;
int g_opt_bwd_depth = 0;
int g_opt_csi_fwd_ppl = 0;   // CSI forward blend: 1 or 2 pixels per lane on the VALU, 0 = the MFMA kernel

// Autotuned blend mappings, per batch shape (gsparc_autotune).
struct TuneKey {
    int P, R, W, H, S;
    bool operator<(const TuneKey& o) const
    {
        return std::tie(P, R, W, H, S) < std::tie(o.P, o.R, o.W, o.H, o.S);
    }
};
struct TuneVal { int fwd_ppl, bwd_ppl, order, fwd_depth, bwd_depth; };
std::mutex g_tune_mu;
std::map<TuneKey, TuneVal> g_tuned;

TuneKey tune_key(const gs::Dims& d) { return {d.Ptot, d.R, d.W, d.H, d.S}; }

bool tuned_for(const gs::Dims& d, TuneVal& v)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(tune_key(d));
    if (it == g_tuned.end()) return false;
    v = it->second;
    return true;
}

int choose_order(const gs::Dims& d)
{
    if (g_opt_order >= 1 && g_opt_order <= 3) return g_opt_order;
    TuneVal v;
    if (tuned_for(d, v)) return v.order;
    // per-XCD receiver sets once every XCD gets at least two receivers
    return d.R >= 16 ? gs::ORDER_XCD_MIXED : gs::ORDER_GLOBAL;
}

int fwd_ppl(const gs::Dims& d)
{
    TuneVal v;
    if (g_opt_fwd_ppl == 0 && tuned_for(d, v)) return v.fwd_ppl;
    return gs::choose_ppl(d, g_opt_fwd_ppl);
}

int bwd_ppl(const gs::Dims& d)
{
    TuneVal v;
    if (g_opt_bwd_ppl == 0 && tuned_for(d, v)) return v.bwd_ppl;
    return gs::choose_ppl(d, g_opt_bwd_ppl);
}

int fwd_depth(const gs::Dims& d)
{
    TuneVal v;
    if (g_opt_fwd_depth == 0 && tuned_for(d, v)) return v.fwd_depth;
    return gs::choose_depth(g_opt_fwd_depth);
}

int bwd_depth(const gs::Dims& d)
{
    TuneVal v;
    if (g_opt_bwd_depth == 0 && tuned_for(d, v)) return v.bwd_depth;
    return gs::choose_depth(g_opt_bwd_depth);
}

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define GS_HIP(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return fail(GSPARC_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr size_t ALIGN = 256;
inline size_t al(size_t x) { return (x + ALIGN - 1) / ALIGN * ALIGN; }

// partial-sum slots of the preprocess backward: one per (receiver group,
// Gaussian of the group's scene)
size_t part_slots(const gs::Dims& d) { return d.Npart; }

void make_layout(const gs::Dims& d, uint64_t K, gsparc_layout& L)
{
    const size_t NR = d.Nrec, N = (size_t)d.W * d.H, RT = (size_t)d.R * d.T;
    const int R = d.R;
    size_t o = 0;
    L.splat = o; o = al(o + NR * 2 * sizeof(float4));
    L.sig = o; o = al(o + NR * sizeof(float2));
    L.depth = o; o = al(o + NR * sizeof(float));
    L.cov3 = o; o = al(o + (size_t)d.Ptot * 6 * sizeof(float));
    L.sort_key0 = o; o = al(o + NR * sizeof(uint32_t));
    L.sort_val0 = o; o = al(o + NR * sizeof(uint32_t));
    L.sort_key1 = o; o = al(o + NR * sizeof(uint32_t));
    L.sort_val1 = o; o = al(o + NR * sizeof(uint32_t));
    L.sort_hist = o; o = al(o + gs::os_words(d) * sizeof(uint32_t));
    L.nvis = o; o = al(o + (size_t)R * sizeof(uint32_t));
    L.plen = o; o = al(o + NR * sizeof(float));
    L.rect = o; o = al(o + NR * sizeof(uint32_t));
    L.srect = o; o = al(o + NR * sizeof(uint32_t));
    L.geom_bytes = o + ALIGN;
    o = 0;
    L.final_t = o; o = al(o + (size_t)R * N * sizeof(float));
    L.n_contrib = o; o = al(o + (size_t)R * N * sizeof(uint32_t));
    L.ranges = o; o = al(o + RT * sizeof(uint2));
    L.segcnt = o; o = al(o + RT * sizeof(uint32_t));
    L.segbase = o; o = al(o + RT * sizeof(uint32_t));
    L.tile_order = o; o = al(o + gs::order_length(d) * sizeof(uint32_t));
    L.chunk_cnt = o; o = al(o + RT * gs::chunk_count_for(d) * sizeof(uint32_t));
    L.total = o; o = al(o + 4 * sizeof(uint32_t));
    L.image_bytes = o + ALIGN;
    o = 0;
    const size_t Kn = (size_t)(K > 0 ? K : 1);
    L.plist = o; o = al(o + Kn * 4);
    L.keys_sorted = o; o = al(o + Kn * 4);
    L.bin_bytes = o + ALIGN;
}

void make_layout(int P, int R, int W, int H, uint64_t K, gsparc_layout& L)
{
    make_layout(gs::make_dims(P, R, W, H), K, L);
}

// grad workspace: g8 [Nrec*8] | dz [Nrec*2] | partials [slots*12] | gL [Nrec]
void grad_layout(const gs::Dims& d, size_t& dz_off, size_t& part_off, size_t& gl_off, size_t& bytes)
{
    dz_off = al(d.Nrec * 8 * sizeof(float));
    part_off = al(dz_off + d.Nrec * 2 * sizeof(float));
    gl_off = al(part_off + part_slots(d) * 12 * sizeof(float));
    bytes = gl_off + d.Nrec * sizeof(float) + ALIGN;
}

// binning state shared by the two forward calls
static gs::BinArgs bin_args(const gs::Dims& d, const gsparc_layout& L, char* gb, char* ib, const int* radii)
{
    gs::BinArgs ba{};
    ba.splat = (const float4*)(gb + L.splat);
    ba.radii = radii;
    ba.depth = (const float*)(gb + L.depth);
    ba.segcnt = (uint32_t*)(ib + L.segcnt);
    ba.segbase = (uint32_t*)(ib + L.segbase);
    ba.ranges = (uint2*)(ib + L.ranges);
    ba.total = (uint32_t*)(ib + L.total);
    ba.nvis = (uint32_t*)(gb + L.nvis);
    ba.skey[0] = (uint32_t*)(gb + L.sort_key0);
    ba.sval[0] = (uint32_t*)(gb + L.sort_val0);
    ba.skey[1] = (uint32_t*)(gb + L.sort_key1);
    ba.sval[1] = (uint32_t*)(gb + L.sort_val1);
    ba.shist = (uint32_t*)(gb + L.sort_hist);
    ba.ccnt = (uint32_t*)(ib + L.chunk_cnt);
    ba.chunk_len = gs::chunk_len_for(d);
    ba.nchunks = gs::chunk_count_for(d);
    if (gs::rect_packed(d)) {
        ba.rect = (const uint32_t*)(gb + L.rect);
        ba.srect = (uint32_t*)(gb + L.srect);
    }
    return ba;
}

// Buffers from torch are 256-byte aligned already; be safe for others.
inline char* base_of(void* p) { return (char*)(((uintptr_t)p + ALIGN - 1) / ALIGN * ALIGN); }

int check_map(int R, int W, int H)
{
    if (R < 0 || W <= 0 || H <= 0) return fail(GSPARC_ERR_ARG, "negative or zero size");
    if ((int64_t)R * W * H > (int64_t)1 << 31) return fail(GSPARC_ERR_ARG, "R*H*W too large");
    if ((W + 15) / 16 + (H + 15) / 16 > 8192) return fail(GSPARC_ERR_ARG, "W + H too large (at most 131072 pixels)");
    // the ordered scatter divides tile indices in float: exact below 2^20 tiles
    if ((int64_t)((W + 15) / 16) * ((H + 15) / 16) >= (int64_t)1 << 20)
        return fail(GSPARC_ERR_ARG, "W * H too large (at most 2^20 tiles of 16x16)");
    return GSPARC_OK;
}

int check_dims(int P, int R, int W, int H)
{
    if (P < 0 || R < 0 || W <= 0 || H <= 0) return fail(GSPARC_ERR_ARG, "negative or zero size");
    if ((int64_t)R * W * H > (int64_t)1 << 31) return fail(GSPARC_ERR_ARG, "R*H*W too large");
    if ((int64_t)P * R >= (int64_t)1 << 31) return fail(GSPARC_ERR_ARG, "P*R too large");
    // the depth sort's look-back words hold a receiver's counts in 28 bits
    if (P >= 1 << 28) return fail(GSPARC_ERR_ARG, "P must be below 2^28");
    // the binning keeps one 64-bit mask per tile column and row in LDS
    if ((W + 15) / 16 + (H + 15) / 16 > 8192) return fail(GSPARC_ERR_ARG, "W + H too large (at most 131072 pixels)");
    // the ordered scatter divides tile indices in float: exact below 2^20 tiles
    if ((int64_t)((W + 15) / 16) * ((H + 15) / 16) >= (int64_t)1 << 20)
        return fail(GSPARC_ERR_ARG, "W * H too large (at most 2^20 tiles of 16x16)");
    return GSPARC_OK;
}

// The device keeps K in 32 bits, saturated at 2^32 - 1 ("that many or more",
// k_binning.hip); a capacity must stay below it so that a saturated count
// always reads as an overflow.
constexpr uint64_t K_SATURATED = 0xffffffffull;

int check_capacity(uint64_t K_capacity)
{
    if (K_capacity >= K_SATURATED) return fail(GSPARC_ERR_ARG, "K_capacity must be below 2^32 - 1");
    return GSPARC_OK;
}

// ---- per-kernel timing (HIP events on the launch stream), off by default
struct ProfPending {
    const char* name;
    hipEvent_t a, b;
};
std::mutex g_prof_mu;
bool g_prof_on = false;      // per-kernel HIP-event timing
bool g_prof_count = false;   // blend kernels' pair / entry counters (their counting variants)
std::vector<ProfPending> g_prof_pending;
std::vector<hipEvent_t> g_prof_free;
std::map<std::string, std::pair<double, long long>> g_prof_acc;
unsigned long long* g_counts_dev = nullptr;   // 4 profiling counters of the blend kernels

hipEvent_t prof_event()
{
    if (!g_prof_free.empty()) {
        hipEvent_t e = g_prof_free.back();
        g_prof_free.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

#define GS_LAUNCH(name, call, stream)                                                   \
    do {                                                                                \
        hipEvent_t a_ = nullptr, b_ = nullptr;                                          \
        if (g_prof_on) {                                                                \
            std::lock_guard<std::mutex> lk_(g_prof_mu);                                 \
            a_ = prof_event();                                                          \
            b_ = prof_event();                                                          \
            if (a_) hipEventRecord(a_, stream);                                         \
        }                                                                               \
        GS_HIP(call);                                                                   \
        if (g_prof_on && a_ && b_) {                                                    \
            std::lock_guard<std::mutex> lk_(g_prof_mu);                                 \
            hipEventRecord(b_, stream);                                                 \
            g_prof_pending.push_back({name, a_, b_});                                   \
        }                                                                               \
    } while (0)

__global__ void k_mark_visible(int P, const float* means, const float* view, uint8_t* present)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    float t[3];
    gs::view_xform(means[3 * i], means[3 * i + 1], means[3 * i + 2], view, t);
    const float rr = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
    present[i] = rr <= 0.04f ? 0 : 1;
}

}  // namespace

extern "C" {

const char* gsparc_version(void) { return "gsparc-mi355x 0.1 (gfx950)"; }
const char* gsparc_last_error(void) { return g_err.c_str(); }

static int* option_slot(const char* name)
{
    if (std::strcmp(name, "render_fwd_ppl") == 0) return &g_opt_fwd_ppl;
    if (std::strcmp(name, "render_bwd_ppl") == 0) return &g_opt_bwd_ppl;
    if (std::strcmp(name, "render_order") == 0) return &g_opt_order;
    if (std::strcmp(name, "csi_fwd_ppl") == 0) return &g_opt_csi_fwd_ppl;
    if (std::strcmp(name, "render_fwd_depth") == 0) return &g_opt_fwd_depth;
    if (std::strcmp(name, "render_bwd_depth") == 0) return &g_opt_bwd_depth;
    return nullptr;
}

int gsparc_set_option(const char* name, int value)
{
    if (!name) return fail(GSPARC_ERR_ARG, "null option name");
    int* slot = option_slot(name);
    if (!slot) return fail(GSPARC_ERR_ARG, std::string("unknown option ") + name);
    if (slot == &g_opt_order) {
        if (value < 0 || value > 3) return fail(GSPARC_ERR_ARG, "render_order must be 0, 1, 2 or 3");
    } else if (slot == &g_opt_csi_fwd_ppl) {
        if (value < 0 || value > 2) return fail(GSPARC_ERR_ARG, "csi_fwd_ppl must be 0, 1 or 2");
    } else if (slot == &g_opt_fwd_depth || slot == &g_opt_bwd_depth) {
        if (value != 0 && value != 32 && value != 64 && value != 128)
            return fail(GSPARC_ERR_ARG, std::string(name) + " must be 0, 32, 64 or 128");
    } else if (value != 0 && value != 1 && value != 2 && value != 4) {
        return fail(GSPARC_ERR_ARG, std::string(name) + " must be 0, 1, 2 or 4");
    }
    *slot = value;
    return GSPARC_OK;
}

int gsparc_get_option(const char* name, int* value)
{
    if (!name || !value) return fail(GSPARC_ERR_ARG, "null pointer");
    int* slot = option_slot(name);
    if (!slot) return fail(GSPARC_ERR_ARG, std::string("unknown option ") + name);
    *value = *slot;
    return GSPARC_OK;
}

size_t gsparc_geom_buffer_size(int P, int R, int W, int H)
{
    gsparc_layout L;
    make_layout(P, R, W, H, 0, L);
    return L.geom_bytes;
}

size_t gsparc_image_buffer_size(int P, int R, int W, int H)
{
    gsparc_layout L;
    make_layout(P, R, W, H, 0, L);
    return L.image_bytes;
}

size_t gsparc_binning_buffer_size(uint64_t K)
{
    gsparc_layout L;
    make_layout(0, 0, 1, 1, K, L);
    return L.bin_bytes;
}

size_t gsparc_grad_buffer_size(int P, int R)
{
    size_t a, b, c, e;
    grad_layout(gs::make_dims(P, R, 1, 1), a, b, c, e);
    return e;
}

size_t gsparc_deterministic_grad_size(size_t grad_bytes, uint64_t K_capacity)
{
    // the per-instance slots, then one u32 per record and per 1024 records:
    // the grad workspace holds 32 bytes per record, so an eighth of it bounds
    // the records' words
    return al(grad_bytes) + al((size_t)(K_capacity > 0 ? K_capacity : 1) * 32) + al(grad_bytes / 8) + 4096;
}

int gsparc_layout_query(int P, int R, int W, int H, uint64_t K, gsparc_layout* out)
{
    if (!out) return fail(GSPARC_ERR_ARG, "null layout");
    make_layout(P, R, W, H, K, *out);
    return GSPARC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- core
// The launch sequences, for any batch shape (uniform or scene table).
namespace {

struct FwdInputs {
    const float *means, *scales, *rotations, *opacities, *mlp, *views, *tx, *rx;
    float mod;
};

int forward_geom_impl(const gs::Dims& d, const FwdInputs& in, void* geom, void* image, int* radii_out,
                      hipStream_t s)
{
    if (!geom || !image || (d.Ptot > 0 && d.R > 0 && (!in.means || !in.scales || !in.rotations ||
                                                      !in.opacities || !in.views || !in.tx || !in.rx ||
                                                      !radii_out)))
        return fail(GSPARC_ERR_ARG, "null pointer");
    if (((uintptr_t)in.mlp & 15) != 0) return fail(GSPARC_ERR_ARG, "emission_mlps must be 16-byte aligned");
    gsparc_layout L;
    make_layout(d, 0, L);
    char* gb = base_of(geom);
    char* ib = base_of(image);
    gs::PreFwdArgs pa;
    pa.means = in.means; pa.scales = in.scales; pa.rots = in.rotations; pa.opac = in.opacities; pa.mlp = in.mlp;
    pa.views = in.views; pa.tx = in.tx; pa.rx = in.rx; pa.mod = in.mod;
    pa.splat = (float4*)(gb + L.splat); pa.sig = (float2*)(gb + L.sig); pa.depth = (float*)(gb + L.depth);
    pa.radii = radii_out; pa.cov3 = (float*)(gb + L.cov3); pa.plen = (float*)(gb + L.plen);
    pa.rect = gs::rect_packed(d) ? (uint32_t*)(gb + L.rect) : nullptr;
    if (d.P > 0 && d.R > 0) GS_LAUNCH("preprocess_fwd", gs::launch_preprocess_fwd(d, pa, s), s);
    if (d.R == 0) return GSPARC_OK;
    gs::BinArgs ba = bin_args(d, L, gb, ib, radii_out);
    if (d.P > 0) {
        GS_LAUNCH("bin_depth_sort", gs::launch_depth_sort(d, ba, s), s);
        GS_LAUNCH("bin_chunk_count", gs::launch_chunk_count(d, ba, s), s);
        GS_LAUNCH("bin_chunk_scan", gs::launch_chunk_scan(d, ba, s), s);
    } else {
        GS_LAUNCH("clear_counts", gs::launch_fill_u32(ba.segcnt, 0u, (size_t)d.R * d.T, s), s);
        // no sort: its wait-limit flag (read by the segment scan) is clear
        GS_LAUNCH("clear_counts", gs::launch_fill_u32(ba.total + 2, 0u, 1, s), s);
    }
    GS_LAUNCH("bin_segscan", gs::launch_segscan_order(d, ba, choose_order(d), (uint32_t*)(ib + L.tile_order), s), s);
    return GSPARC_OK;
}

int read_count_impl(const gs::Dims& d, const void* image, uint64_t* K, hipStream_t s)
{
    *K = 0;
    if (d.R == 0) return GSPARC_OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    GS_HIP(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone)
        return fail(GSPARC_ERR_ARG, "the instance count synchronises the stream: not allowed while capturing; "
                                    "use gsparc_rasterize_forward with a capacity");
    gsparc_layout L;
    make_layout(d, 0, L);
    uint32_t tot = 0;
    GS_HIP(hipMemcpyAsync(&tot, base_of((void*)image) + L.total, sizeof(tot), hipMemcpyDeviceToHost, s));
    GS_HIP(hipStreamSynchronize(s));
    *K = tot;
    return GSPARC_OK;
}

int forward_render_impl(const gs::Dims& d, const float* emission, void* geom, void* image, void* bin,
                        uint64_t K_capacity, const int* radii, float* out_real, float* out_imag, int flags,
                        hipStream_t s)
{
    if (!geom || !image || !bin || (d.R > 0 && (!out_real || !out_imag || (d.P > 0 && !radii))))
        return fail(GSPARC_ERR_ARG, "null pointer");
    if (int e = check_capacity(K_capacity)) return e;
    if (d.R == 0) return GSPARC_OK;
    gsparc_layout L;
    make_layout(d, K_capacity, L);
    char* gb = base_of(geom);
    char* ib = base_of(image);
    char* bb = base_of(bin);
    gs::BinArgs ba = bin_args(d, L, gb, ib, radii);
    ba.plist = (uint32_t*)(bb + L.plist);
    ba.keys_sorted = (flags & GSPARC_FLAG_KEEP_KEYS) ? (uint32_t*)(bb + L.keys_sorted) : nullptr;
    ba.Kcap = K_capacity;
    if (d.P > 0) GS_LAUNCH("bin_scatter", gs::launch_bin_place(d, ba, s), s);
    gs::RenderFwdArgs ra;
    ra.ranges = ba.ranges;
    ra.order = (const uint32_t*)(ib + L.tile_order);
    ra.plist = ba.plist;
    ra.splat = ba.splat;
    ra.sig = emission ? (const float2*)emission : (const float2*)(gb + L.sig);
    ra.out_re = out_real;
    ra.out_im = out_imag;
    ra.finalT = (float*)(ib + L.final_t);
    ra.ncontrib = (uint32_t*)(ib + L.n_contrib);
    ra.total = ba.total;
    ra.Kcap = K_capacity;
    ra.counts = g_prof_count ? g_counts_dev : nullptr;
    GS_LAUNCH("render_fwd", gs::launch_render_fwd(d, ra, fwd_ppl(d), fwd_depth(d), s), s);
    return GSPARC_OK;
}

struct BwdOutputs {
    float *dmeans2D, *dmeans3D, *dmlp, *dopacity, *dcov3D, *dscales, *drotations, *demission;
};

// csi != nullptr: the multi-frequency blend's adjoint (exact mode), with
// dout_real the [R*N*NC*2] gradient of its output (dout_imag unused).
int backward_impl(int grad_mode, const gs::Dims& d, const FwdInputs& in, const float* emission, const int* radii,
                  void* geom, void* image, void* bin, uint64_t K_capacity, void* grad, const float* dout_real,
                  const float* dout_imag, const BwdOutputs& out, hipStream_t s, const GsCsi* csi = nullptr)
{
    const bool det = (grad_mode & GSPARC_GRAD_DETERMINISTIC) != 0;
    grad_mode &= ~GSPARC_GRAD_DETERMINISTIC;
    if (grad_mode != GSPARC_GRAD_EXACT && grad_mode != GSPARC_GRAD_REFERENCE)
        return fail(GSPARC_ERR_ARG, "unknown grad_mode");
    if (det && csi) return fail(GSPARC_ERR_ARG, "deterministic gradients: single-channel backward only");
    if (d.P == 0 || d.R == 0) return GSPARC_OK;
    if (!geom || !image || !bin || !grad || !in.means || !in.scales || !in.rotations || !in.views || !in.tx ||
        !in.rx || !radii || !dout_real || (!csi && !dout_imag) || !out.dmeans3D || !out.dopacity ||
        !out.dscales || !out.drotations)
        return fail(GSPARC_ERR_ARG, "null pointer");
    if (((uintptr_t)in.mlp & 15) != 0 || ((uintptr_t)out.dmlp & 15) != 0)
        return fail(GSPARC_ERR_ARG, "emission_mlps and its gradient must be 16-byte aligned");
    if (int e = check_capacity(K_capacity)) return e;
    gsparc_layout L;
    make_layout(d, K_capacity, L);
    char* gb = base_of(geom);
    char* ib = base_of(image);
    char* bb = base_of(bin);
    float* g8 = (float*)base_of(grad);
    size_t dz_off, part_off, gl_off, bytes;
    grad_layout(d, dz_off, part_off, gl_off, bytes);
    float* gL = (float*)((char*)g8 + gl_off);
    // the accumulators start at zero (the deterministic gather writes them all)
    if (!det) GS_LAUNCH("grad_clear", gs::launch_fill_u32((uint32_t*)g8, 0u, d.Nrec * 8, s), s);
    if (csi) {
        GS_LAUNCH("grad_clear", gs::launch_fill_u32((uint32_t*)gL, 0u, d.Nrec, s), s);
        gs::RenderCsiArgs ca;
        ca.ranges = (const uint2*)(ib + L.ranges);
        ca.order = (const uint32_t*)(ib + L.tile_order);
        ca.plist = (const uint32_t*)(bb + L.plist);
        ca.splat = (const float4*)(gb + L.splat);
        ca.sig = emission ? (const float2*)emission : (const float2*)(gb + L.sig);
        ca.plen = (const float*)(gb + L.plen);
        ca.out = nullptr;
        ca.finalT = (float*)(ib + L.final_t);
        ca.ncontrib = (uint32_t*)(ib + L.n_contrib);
        ca.dout = dout_real;
        ca.g8 = g8;
        ca.gL = gL;
        ca.total = (const uint32_t*)(ib + L.total);
        ca.Kcap = K_capacity;
        ca.csi = *csi;
        GS_LAUNCH("render_bwd_csi", gs::launch_render_bwd_csi(d, ca, s), s);
    } else {
        gs::RenderBwdArgs ra;
        ra.ranges = (const uint2*)(ib + L.ranges);
        ra.order = (const uint32_t*)(ib + L.tile_order);
        ra.plist = (const uint32_t*)(bb + L.plist);
        ra.splat = (const float4*)(gb + L.splat);
        ra.sig = emission ? (const float2*)emission : (const float2*)(gb + L.sig);
        ra.finalT = (const float*)(ib + L.final_t);
        ra.ncontrib = (const uint32_t*)(ib + L.n_contrib);
        ra.dout_re = dout_real;
        ra.dout_im = dout_imag;
        ra.g8 = g8;
        ra.mode = grad_mode;
        ra.total = (const uint32_t*)(ib + L.total);
        ra.Kcap = K_capacity;
        ra.counts = g_prof_count ? g_counts_dev : nullptr;
        ra.inst8 = nullptr;
        ra.det_off = nullptr;
        ra.radii = radii;
        gs::DetArgs da;
        if (det) {
            // after the grad workspace: the per-instance slots, then each
            // record's first slot and the scan's block sums
            ra.inst8 = (float*)((char*)g8 + al(bytes));
            uint32_t* off = (uint32_t*)((char*)ra.inst8 + al((size_t)std::max<uint64_t>(K_capacity, 1) * 32));
            da.splat = ra.splat;
            da.radii = radii;
            da.off = off;
            da.bsum = off + d.Nrec;
            da.inst8 = ra.inst8;
            da.g8 = g8;
            da.total = ra.total;
            da.Kcap = K_capacity;
            ra.det_off = off;
            ra.counts = nullptr;
            GS_LAUNCH("det_clear", gs::launch_fill_u32((uint32_t*)ra.inst8, 0u, (size_t)K_capacity * 8, s), s);
            GS_LAUNCH("det_offsets", gs::launch_det_offsets(d, da, s), s);
        }
        GS_LAUNCH("render_bwd", gs::launch_render_bwd(d, ra, bwd_ppl(d), bwd_depth(d), s), s);
        if (det) GS_LAUNCH("det_gather", gs::launch_det_gather(d, da, s), s);
    }
    gs::PreBwdArgs pa;
    pa.means = in.means; pa.scales = in.scales; pa.rots = in.rotations; pa.mlp = in.mlp; pa.views = in.views;
    pa.sig = (const float2*)(gb + L.sig); pa.cov3 = (const float*)(gb + L.cov3);
    pa.splat = (const float4*)(gb + L.splat);
    pa.tx = in.tx; pa.rx = in.rx; pa.mod = in.mod; pa.radii = radii; pa.g8 = g8;
    pa.dmeans = out.dmeans3D; pa.dscales = out.dscales; pa.drots = out.drotations; pa.dopac = out.dopacity;
    pa.dmlp = out.dmlp; pa.dcov3 = out.dcov3D; pa.dmeans2d = out.dmeans2D; pa.dsig = out.demission;
    pa.gL = csi ? gL : nullptr;
    pa.mode = grad_mode;
    pa.dz = (float2*)((char*)g8 + dz_off);
    pa.part = (float*)((char*)g8 + part_off);
    GS_LAUNCH("preprocess_bwd", gs::launch_preprocess_bwd(d, pa, s), s);
    return GSPARC_OK;
}

// Multi-frequency blend forward on the geometry of a finished forward_geom:
// ordered scatter into the tile lists, then the CSI blend.
int forward_render_csi_impl(const gs::Dims& d, const GsCsi& csi, const float* emission, void* geom,
                            void* image, void* bin, uint64_t K_capacity, const int* radii, float* out,
                            hipStream_t s)
{
    if (!geom || !image || !bin || (d.R > 0 && (!out || (d.P > 0 && !radii))))
        return fail(GSPARC_ERR_ARG, "null pointer");
    if (int e = check_capacity(K_capacity)) return e;
    if (d.R == 0) return GSPARC_OK;
    gsparc_layout L;
    make_layout(d, K_capacity, L);
    char* gb = base_of(geom);
    char* ib = base_of(image);
    char* bb = base_of(bin);
    gs::BinArgs ba = bin_args(d, L, gb, ib, radii);
    ba.plist = (uint32_t*)(bb + L.plist);
    ba.keys_sorted = nullptr;
    ba.Kcap = K_capacity;
    if (d.P > 0) GS_LAUNCH("bin_scatter", gs::launch_bin_place(d, ba, s), s);
    gs::RenderCsiArgs ca;
    ca.ranges = ba.ranges;
    ca.order = (const uint32_t*)(ib + L.tile_order);
    ca.plist = ba.plist;
    ca.splat = ba.splat;
    ca.sig = emission ? (const float2*)emission : (const float2*)(gb + L.sig);
    ca.plen = (const float*)(gb + L.plen);
    ca.out = out;
    ca.finalT = (float*)(ib + L.final_t);
    ca.ncontrib = (uint32_t*)(ib + L.n_contrib);
    ca.dout = nullptr; ca.g8 = nullptr; ca.gL = nullptr;
    ca.total = ba.total;
    ca.Kcap = K_capacity;
    ca.csi = csi;
    // 0: the matrix-core forward; 1, 2: the VALU forward with that many pixels per lane
    GS_LAUNCH("render_fwd_csi", gs::launch_render_fwd_csi(d, ca, g_opt_csi_fwd_ppl, s), s);
    return GSPARC_OK;
}

int check_csi(double f0_hz, double df_hz, GsCsi& t)
{
    if (!(f0_hz >= 0.0) || !(df_hz >= 0.0) || !(f0_hz + GS_CSI_NC * df_hz < 1e12))
        return fail(GSPARC_ERR_ARG, "subcarrier frequencies must be finite, non-negative and below 1 THz");
    t = gs_csi_make(f0_hz, df_hz, GS_LIGHT_SPEED);
    return GSPARC_OK;
}

// Times every blend mapping on the buffers of a finished forward and keeps
// the fastest per batch shape.  Each timed launch runs where it runs in a
// training step: the forward blend right after the ordered scatter that
// writes its tile lists, the backward blend right after a forward blend and
// the gradient clear, so caches hold what they hold in the step (back-to-back
// launches of one kernel would keep its lists hot and favour the mappings
// that re-read them).  The forward result does not depend on the mapping
// (bit-identical outputs); the backward only by summation order.
int autotune_impl(int grad_mode, const gs::Dims& d, const float* emission, const int* radii, void* geom,
                  void* image, void* bin, uint64_t K_capacity, void* grad, const float* dout_real,
                  const float* dout_imag, float* out_real, float* out_imag, int iters, int* choice, float* ms,
                  float* table, hipStream_t s)
{
    if (grad_mode != GSPARC_GRAD_EXACT && grad_mode != GSPARC_GRAD_REFERENCE)
        return fail(GSPARC_ERR_ARG, "unknown grad_mode");
    if (!geom || !image || !bin || !grad || !radii || !dout_real || !dout_imag || !out_real || !out_imag ||
        !choice)
        return fail(GSPARC_ERR_ARG, "null pointer");
    if (d.R == 0 || d.Ptot == 0) return fail(GSPARC_ERR_ARG, "nothing to tune");
    if (int e = check_capacity(K_capacity)) return e;
    if (g_prof_on || g_prof_count) return fail(GSPARC_ERR_ARG, "disable profiling before tuning");
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    GS_HIP(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone) return fail(GSPARC_ERR_ARG, "tuning times kernels: not while capturing");
    iters = std::max(1, std::min(iters, 100));
    gsparc_layout L;
    make_layout(d, K_capacity, L);
    char* gb = base_of(geom);
    char* ib = base_of(image);
    char* bb = base_of(bin);
    float* g8 = (float*)base_of(grad);
    gs::BinArgs sa = bin_args(d, L, gb, ib, radii);
    sa.plist = (uint32_t*)(bb + L.plist);
    sa.keys_sorted = nullptr;
    sa.Kcap = K_capacity;
    uint32_t* order = (uint32_t*)(ib + L.tile_order);
    gs::RenderFwdArgs fa;
    fa.ranges = sa.ranges; fa.order = order; fa.plist = sa.plist; fa.splat = sa.splat;
    fa.sig = emission ? (const float2*)emission : (const float2*)(gb + L.sig);
    fa.out_re = out_real; fa.out_im = out_imag;
    fa.finalT = (float*)(ib + L.final_t); fa.ncontrib = (uint32_t*)(ib + L.n_contrib);
    fa.total = sa.total; fa.Kcap = K_capacity; fa.counts = nullptr;
    gs::RenderBwdArgs ba;
    ba.ranges = sa.ranges; ba.order = order; ba.plist = sa.plist; ba.splat = sa.splat; ba.sig = fa.sig;
    ba.finalT = fa.finalT; ba.ncontrib = fa.ncontrib; ba.dout_re = dout_real; ba.dout_im = dout_imag; ba.g8 = g8;
    ba.mode = grad_mode; ba.total = sa.total; ba.Kcap = K_capacity; ba.counts = nullptr; ba.inst8 = nullptr;
    ba.det_off = nullptr; ba.radii = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    GS_HIP(hipEventCreate(&ev[0]));
    if (hipEventCreate(&ev[1]) != hipSuccess) {
        hipEventDestroy(ev[0]);
        return fail(GSPARC_ERR_HIP, "hipEventCreate");
    }
    // median of `iters` single timed launches, each after its predecessors
    auto time_it = [&](auto&& before, auto&& launch, float& out_ms) -> hipError_t {
        std::vector<float> t(iters + 1);
        hipError_t e = hipSuccess;
        for (int i = 0; i <= iters && e == hipSuccess; ++i) {   // launch 0 warms up
            e = before();
            if (e == hipSuccess) e = hipEventRecord(ev[0], s);
            if (e == hipSuccess) e = launch();
            if (e == hipSuccess) e = hipEventRecord(ev[1], s);
            if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
            if (e == hipSuccess) e = hipEventElapsedTime(&t[i], ev[0], ev[1]);
        }
        std::sort(t.begin() + 1, t.end());
        out_ms = t[1 + iters / 2];
        return e;
    };
    // candidates: pixels per lane x batch depth, forward and backward
    const int ppls[3] = {1, 2, 4};
    const int depths[3] = {32, 64, 128};
    TuneVal best{1, 1, 1, 64, 64};
    float best_f = 0.0f, best_b = 0.0f, best_total = 1e30f;
    hipError_t err = hipSuccess;
    auto scatter = [&] { return gs::launch_bin_place(d, sa, s); };
    for (int mode = 1; mode <= 3 && err == hipSuccess; ++mode) {
        err = gs::launch_tile_order(d, mode, sa.ranges, order, sa.total, s);
        float bf = 1e30f, bbk = 1e30f;
        int pf = 1, pb = 1, df = 64, db = 64;
        for (int k = 0; k < 9 && err == hipSuccess; ++k) {
            const int ppl = ppls[k / 3], dep = depths[k % 3];
            float t = 0.0f;
            err = time_it(scatter, [&] { return gs::launch_render_fwd(d, fa, ppl, dep, s); }, t);
            if (table) table[(mode - 1) * 18 + k] = t;
            if (t < bf) { bf = t; pf = ppl; df = dep; }
        }
        for (int k = 0; k < 9 && err == hipSuccess; ++k) {
            const int ppl = ppls[k / 3], dep = depths[k % 3];
            float t = 0.0f;
            auto before = [&] {
                hipError_t e = scatter();
                if (e == hipSuccess) e = gs::launch_render_fwd(d, fa, pf, df, s);
                if (e == hipSuccess) e = gs::launch_fill_u32((uint32_t*)g8, 0u, d.Nrec * 8, s);
                return e;
            };
            err = time_it(before, [&] { return gs::launch_render_bwd(d, ba, ppl, dep, s); }, t);
            if (table) table[(mode - 1) * 18 + 9 + k] = t;
            if (t < bbk) { bbk = t; pb = ppl; db = dep; }
        }
        if (err == hipSuccess && bf + bbk < best_total) {
            best_total = bf + bbk;
            best = {pf, pb, mode, df, db};
            best_f = bf;
            best_b = bbk;
        }
    }
    // leave the buffers as the chosen mapping's forward would
    if (err == hipSuccess) err = gs::launch_tile_order(d, best.order, sa.ranges, order, sa.total, s);
    if (err == hipSuccess) err = scatter();
    if (err == hipSuccess) err = gs::launch_render_fwd(d, fa, best.fwd_ppl, best.fwd_depth, s);
    if (err == hipSuccess) err = hipStreamSynchronize(s);
    hipEventDestroy(ev[0]);
    hipEventDestroy(ev[1]);
    if (err != hipSuccess) return fail(GSPARC_ERR_HIP, std::string("autotune: ") + hipGetErrorString(err));
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[tune_key(d)] = best;
    }
    choice[0] = best.fwd_ppl;
    choice[1] = best.bwd_ppl;
    choice[2] = best.order;
    choice[3] = best.fwd_depth;
    choice[4] = best.bwd_depth;
    if (ms) {
        ms[0] = best_f;
        ms[1] = best_b;
    }
    return GSPARC_OK;
}

// Scene-table batch: the Dims of a batch built by gsparc_scene_table_build.
int scene_dims(const gsparc_batch* b, const uint32_t* table, int W, int H, gs::Dims& d)
{
    if (!b || !table) return fail(GSPARC_ERR_ARG, "null batch or scene table");
    if (b->S < 1 || b->R < 0 || b->G < 0 || b->Pmax < 0 || b->Ptot < 0) return fail(GSPARC_ERR_ARG, "bad batch");
    if (b->Nrec >= ((uint64_t)1 << 31)) return fail(GSPARC_ERR_ARG, "too many (receiver, Gaussian) records");
    // the depth sort's look-back words hold a receiver's counts in 28 bits
    if (b->Pmax >= 1 << 28) return fail(GSPARC_ERR_ARG, "a scene's Gaussian count must be below 2^28");
    if (int e = check_map(b->R, W, H)) return e;
    d = gs::make_dims(b->Pmax, b->R, W, H);
    d.Ptot = b->Ptot;
    d.G = b->G;
    d.S = b->S;
    d.Nrec = (size_t)b->Nrec;
    d.Npart = (size_t)b->Npart;
    d.tab = table;
    return GSPARC_OK;
}

}  // namespace

extern "C" {

int gsparc_rasterize_forward_geom(int P, int R, int W, int H, const float* means, const float* scales,
                                  const float* rotations, const float* opacities, const float* emission_mlps,
                                  const float* viewmatrices, const float* tx_pos, const float* rx_pos,
                                  float scale_modifier, void* geom, void* image, int* radii_out,
                                  uint64_t* K_out, void* stream)
{
    if (int e = check_dims(P, R, W, H)) return e;
    if (!K_out) return fail(GSPARC_ERR_ARG, "null pointer");
    const gs::Dims d = gs::make_dims(P, R, W, H);
    const FwdInputs in{means, scales, rotations, opacities, emission_mlps, viewmatrices, tx_pos, rx_pos,
                       scale_modifier};
    hipStream_t s = (hipStream_t)stream;
    if (int e = forward_geom_impl(d, in, geom, image, radii_out, s)) return e;
    if (int e = read_count_impl(d, image, K_out, s)) return e;
    if (*K_out >= K_SATURATED)
        return fail(GSPARC_ERR_CAPACITY, "the tile lists would hold 2^32 - 1 or more instances");
    return GSPARC_OK;
}

int gsparc_instance_count(int P, int R, int W, int H, const void* image, uint64_t* K, void* stream)
{
    if (int e = check_dims(P, R, W, H)) return e;
    if (!image || !K) return fail(GSPARC_ERR_ARG, "null pointer");
    return read_count_impl(gs::make_dims(P, R, W, H), image, K, (hipStream_t)stream);
}

int gsparc_rasterize_forward_render(int P, int R, int W, int H, const float* emission, void* geom, void* image,
                                    void* bin, uint64_t K_capacity, const int* radii, float* out_real,
                                    float* out_imag, int flags, void* stream)
{
    if (int e = check_dims(P, R, W, H)) return e;
    return forward_render_impl(gs::make_dims(P, R, W, H), emission, geom, image, bin, K_capacity, radii, out_real,
                               out_imag, flags, (hipStream_t)stream);
}

int gsparc_rasterize_forward(int P, int R, int W, int H, const float* means, const float* scales,
                             const float* rotations, const float* opacities, const float* emission_mlps,
                             const float* emission, const float* viewmatrices, const float* tx_pos,
                             const float* rx_pos, float scale_modifier, void* geom, void* image, void* bin,
                             uint64_t K_capacity, int* radii_out, float* out_real, float* out_imag, int flags,
                             void* stream)
{
    if (int e = check_dims(P, R, W, H)) return e;
    const gs::Dims d = gs::make_dims(P, R, W, H);
    const FwdInputs in{means, scales, rotations, opacities, emission_mlps, viewmatrices, tx_pos, rx_pos,
                       scale_modifier};
    hipStream_t s = (hipStream_t)stream;
    if (int e = forward_geom_impl(d, in, geom, image, radii_out, s)) return e;
    return forward_render_impl(d, emission, geom, image, bin, K_capacity, radii_out, out_real, out_imag, flags, s);
}

int gsparc_rasterize_backward(int grad_mode, int P, int R, int W, int H, const float* means, const float* scales,
                              const float* rotations, const float* emission_mlps, const float* emission,
                              const float* viewmatrices, const float* tx_pos, const float* rx_pos,
                              float scale_modifier, const int* radii, void* geom, void* image, void* bin,
                              uint64_t K_capacity, void* grad, const float* dout_real, const float* dout_imag, float* dmeans2D,
                              float* dmeans3D, float* demission_mlps, float* dopacity, float* dcov3D,
                              float* dscales, float* drotations, float* demission, void* stream)
{
    if (int e = check_dims(P, R, W, H)) return e;
    const FwdInputs in{means, scales, rotations, nullptr, emission_mlps, viewmatrices, tx_pos, rx_pos,
                       scale_modifier};
    const BwdOutputs out{dmeans2D, dmeans3D, demission_mlps, dopacity, dcov3D, dscales, drotations, demission};
    return backward_impl(grad_mode, gs::make_dims(P, R, W, H), in, emission, radii, geom, image, bin, K_capacity,
                         grad, dout_real, dout_imag, out, (hipStream_t)stream);
}

// ---------------------------------------------------------------- scene batches
size_t gsparc_scene_table_words(int S, int R)
{
    return (size_t)GS_TAB_HDR + 8 * (size_t)(R > 0 ? R : 0) + 8 * (size_t)(S > 0 ? S : 0);
}

int gsparc_scene_table_build(int S, const int* scene_P, const int* scene_R, uint32_t* table, size_t table_words,
                             gsparc_batch* batch)
{
    if (S < 1 || !scene_P || !scene_R || !table || !batch) return fail(GSPARC_ERR_ARG, "bad scene description");
    int64_t R = 0, Ptot = 0, G = 0, Pmax = 0, Nrec = 0, Npart = 0;
    for (int s = 0; s < S; ++s) {
        if (scene_P[s] < 0 || scene_R[s] < 0) return fail(GSPARC_ERR_ARG, "negative scene size");
        R += scene_R[s];
        Ptot += scene_P[s];
        G += (scene_R[s] + GS_RC - 1) / GS_RC;
        Pmax = std::max<int64_t>(Pmax, scene_P[s]);
        Nrec += (int64_t)scene_R[s] * scene_P[s];
        Npart += (int64_t)((scene_R[s] + GS_RC - 1) / GS_RC) * scene_P[s];
    }
    if (Nrec >= ((int64_t)1 << 31) || Ptot >= ((int64_t)1 << 31) || Npart >= ((int64_t)1 << 32))
        return fail(GSPARC_ERR_ARG, "batch too large");
    if (table_words < gsparc_scene_table_words(S, (int)R)) return fail(GSPARC_ERR_ARG, "scene table too small");
    uint32_t* hdr = table;
    hdr[0] = (uint32_t)S; hdr[1] = (uint32_t)R; hdr[2] = (uint32_t)G; hdr[3] = (uint32_t)Pmax;
    hdr[4] = (uint32_t)Ptot; hdr[5] = (uint32_t)Nrec; hdr[6] = (uint32_t)Npart; hdr[7] = 0;
    uint32_t* rx = table + GS_TAB_HDR;
    uint32_t* gr = rx + 4 * R;
    uint32_t* sc = gr + 4 * G;
    uint32_t roff = 0, g0 = 0, r0 = 0, y = 0, part = 0;
    for (int s = 0; s < S; ++s) {
        const uint32_t P = (uint32_t)scene_P[s], Rs = (uint32_t)scene_R[s];
        const uint32_t ng = (Rs + GS_RC - 1) / GS_RC;
        uint32_t* e = sc + 8 * s;
        e[0] = g0; e[1] = P; e[2] = r0; e[3] = Rs; e[4] = y; e[5] = ng; e[6] = 0; e[7] = 0;
        for (uint32_t i = 0; i < Rs; ++i) {
            uint32_t* f = rx + 4 * (r0 + i);
            f[0] = roff; f[1] = g0; f[2] = P; f[3] = (uint32_t)s;
            roff += P;
        }
        for (uint32_t k = 0; k < ng; ++k) {
            uint32_t* f = gr + 4 * (y + k);
            f[0] = r0 + k * GS_RC; f[1] = std::min<uint32_t>(GS_RC, Rs - k * GS_RC); f[2] = (uint32_t)s; f[3] = part;
            part += P;
        }
        g0 += P;
        r0 += Rs;
        y += ng;
    }
    batch->S = S; batch->R = (int)R; batch->G = (int)G; batch->Pmax = (int)Pmax; batch->Ptot = (int)Ptot;
    batch->Nrec = (uint64_t)Nrec;
    batch->Npart = (uint64_t)Npart;
    return GSPARC_OK;
}

int gsparc_scene_buffer_sizes(const gsparc_batch* batch, int W, int H, uint64_t K, size_t* sizes)
{
    if (!batch || !sizes) return fail(GSPARC_ERR_ARG, "null pointer");
    gs::Dims d = gs::make_dims(batch->Pmax, batch->R, W, H);
    d.Ptot = batch->Ptot; d.G = batch->G; d.S = batch->S; d.Nrec = (size_t)batch->Nrec;
    d.Npart = (size_t)batch->Npart;
    gsparc_layout L;
    make_layout(d, K, L);
    size_t a, b, c, e;
    grad_layout(d, a, b, c, e);
    sizes[0] = L.geom_bytes;
    sizes[1] = L.image_bytes;
    sizes[2] = L.bin_bytes;
    sizes[3] = e;
    return GSPARC_OK;
}

int gsparc_rasterize_forward_scenes(const gsparc_batch* batch, const uint32_t* scene_table, int W, int H,
                                    const float* means, const float* scales, const float* rotations,
                                    const float* opacities, const float* emission_mlps, const float* emission,
                                    const float* viewmatrices, const float* tx_pos, const float* rx_pos,
                                    float scale_modifier, void* geom, void* image, void* bin, uint64_t K_capacity,
                                    int* radii_out, float* out_real, float* out_imag, int flags, void* stream)
{
    gs::Dims d;
    if (int e = scene_dims(batch, scene_table, W, H, d)) return e;
    const FwdInputs in{means, scales, rotations, opacities, emission_mlps, viewmatrices, tx_pos, rx_pos,
                       scale_modifier};
    hipStream_t s = (hipStream_t)stream;
    if (int e = forward_geom_impl(d, in, geom, image, radii_out, s)) return e;
    return forward_render_impl(d, emission, geom, image, bin, K_capacity, radii_out, out_real, out_imag, flags, s);
}

int gsparc_rasterize_forward_geom_scenes(const gsparc_batch* batch, const uint32_t* scene_table, int W, int H,
                                         const float* means, const float* scales, const float* rotations,
                                         const float* opacities, const float* emission_mlps,
                                         const float* viewmatrices, const float* tx_pos, const float* rx_pos,
                                         float scale_modifier, void* geom, void* image, int* radii_out,
                                         uint64_t* K_out, void* stream)
{
    gs::Dims d;
    if (int e = scene_dims(batch, scene_table, W, H, d)) return e;
    if (!K_out) return fail(GSPARC_ERR_ARG, "null pointer");
    const FwdInputs in{means, scales, rotations, opacities, emission_mlps, viewmatrices, tx_pos, rx_pos,
                       scale_modifier};
    hipStream_t s = (hipStream_t)stream;
    if (int e = forward_geom_impl(d, in, geom, image, radii_out, s)) return e;
    if (int e = read_count_impl(d, image, K_out, s)) return e;
    if (*K_out >= K_SATURATED)
        return fail(GSPARC_ERR_CAPACITY, "the tile lists would hold 2^32 - 1 or more instances");
    return GSPARC_OK;
}

int gsparc_rasterize_forward_render_scenes(const gsparc_batch* batch, const uint32_t* scene_table, int W, int H,
                                           const float* emission, void* geom, void* image, void* bin,
                                           uint64_t K_capacity, const int* radii, float* out_real,
                                           float* out_imag, int flags, void* stream)
{
    gs::Dims d;
    if (int e = scene_dims(batch, scene_table, W, H, d)) return e;
    return forward_render_impl(d, emission, geom, image, bin, K_capacity, radii, out_real, out_imag, flags,
                               (hipStream_t)stream);
}

int gsparc_instance_count_scenes(const gsparc_batch* batch, const uint32_t* scene_table, int W, int H,
                                 const void* image, uint64_t* K, void* stream)
{
    gs::Dims d;
    if (int e = scene_dims(batch, scene_table, W, H, d)) return e;
    if (!image || !K) return fail(GSPARC_ERR_ARG, "null pointer");
    return read_count_impl(d, image, K, (hipStream_t)stream);
}

int gsparc_rasterize_backward_scenes(int grad_mode, const gsparc_batch* batch, const uint32_t* scene_table, int W,
                                     int H, const float* means, const float* scales, const float* rotations,
                                     const float* emission_mlps, const float* emission, const float* viewmatrices,
                                     const float* tx_pos, const float* rx_pos, float scale_modifier,
                                     const int* radii, void* geom, void* image, void* bin, uint64_t K_capacity,
                                     void* grad, const float* dout_real, const float* dout_imag, float* dmeans2D,
                                     float* dmeans3D, float* demission_mlps, float* dopacity, float* dcov3D,
                                     float* dscales, float* drotations, float* demission, void* stream)
{
    gs::Dims d;
    if (int e = scene_dims(batch, scene_table, W, H, d)) return e;
    const FwdInputs in{means, scales, rotations, nullptr, emission_mlps, viewmatrices, tx_pos, rx_pos,
                       scale_modifier};
    const BwdOutputs out{dmeans2D, dmeans3D, demission_mlps, dopacity, dcov3D, dscales, drotations, demission};
    return backward_impl(grad_mode, d, in, emission, radii, geom, image, bin, K_capacity, grad, dout_real,
                         dout_imag, out, (hipStream_t)stream);
}

int gsparc_mark_visible(int P, const float* means, const float* viewmatrix, uint8_t* present, void* stream)
{
    if (P < 0) return fail(GSPARC_ERR_ARG, "negative size");
    if (P == 0) return GSPARC_OK;
    if (!means || !viewmatrix || !present) return fail(GSPARC_ERR_ARG, "null pointer");
    hipLaunchKernelGGL(k_mark_visible, dim3((P + 255) / 256), dim3(256), 0, (hipStream_t)stream, P, means,
                       viewmatrix, present);
    GS_HIP(hipGetLastError());
    return GSPARC_OK;
}

int gsparc_ssim_forward(int planes, int H, int W, float C1, float C2, const float* img1, const float* img2,
                        float* ssim_map, float* dm_dmu1, float* dm_dsigma1_sq, float* dm_dsigma12, void* stream)
{
    if (planes < 0 || H < 0 || W < 0) return fail(GSPARC_ERR_ARG, "negative size");
    if (planes > 65535) return fail(GSPARC_ERR_ARG, "at most 65535 planes per call");
    if ((size_t)planes * H * W == 0) return GSPARC_OK;
    if (!img1 || !img2 || !ssim_map) return fail(GSPARC_ERR_ARG, "null pointer");
    if (dm_dmu1 && (!dm_dsigma1_sq || !dm_dsigma12)) return fail(GSPARC_ERR_ARG, "partial derivative maps");
    gs::SsimArgs a{planes, H, W, C1, C2, img1, img2, ssim_map, dm_dmu1, dm_dsigma1_sq, dm_dsigma12, nullptr};
    GS_LAUNCH("ssim_fwd", gs::launch_ssim_fwd(a, (hipStream_t)stream), (hipStream_t)stream);
    return GSPARC_OK;
}

int gsparc_ssim_backward(int planes, int H, int W, float C1, float C2, const float* img1, const float* img2,
                         const float* dL_dmap, const float* dm_dmu1, const float* dm_dsigma1_sq,
                         const float* dm_dsigma12, float* dL_dimg1, void* stream)
{
    (void)C1; (void)C2;
    if (planes < 0 || H < 0 || W < 0) return fail(GSPARC_ERR_ARG, "negative size");
    if (planes > 65535) return fail(GSPARC_ERR_ARG, "at most 65535 planes per call");
    if ((size_t)planes * H * W == 0) return GSPARC_OK;
    if (!img1 || !img2 || !dL_dmap || !dm_dmu1 || !dm_dsigma1_sq || !dm_dsigma12 || !dL_dimg1)
        return fail(GSPARC_ERR_ARG, "null pointer");
    gs::SsimBwdArgs a{planes, H, W, img1, img2, dL_dmap, dm_dmu1, dm_dsigma1_sq, dm_dsigma12, dL_dimg1, 0.0f, 0.0f};
    GS_LAUNCH("ssim_bwd", gs::launch_ssim_bwd(a, (hipStream_t)stream), (hipStream_t)stream);
    return GSPARC_OK;
}

int gsparc_rasterize_forward_render_csi(int P, int R, int W, int H, double f0_hz, double df_hz,
                                        const float* emission, void* geom, void* image, void* bin,
                                        uint64_t K_capacity, const int* radii, float* out, void* stream)
{
    if (int e = check_dims(P, R, W, H)) return e;
    GsCsi t;
    if (int e = check_csi(f0_hz, df_hz, t)) return e;
    return forward_render_csi_impl(gs::make_dims(P, R, W, H), t, emission, geom, image, bin, K_capacity, radii, out,
                                   (hipStream_t)stream);
}

int gsparc_rasterize_forward_csi(int P, int R, int W, int H, double f0_hz, double df_hz, const float* means,
                                 const float* scales, const float* rotations, const float* opacities,
                                 const float* emission_mlps, const float* emission, const float* viewmatrices,
                                 const float* tx_pos, const float* rx_pos, float scale_modifier, void* geom,
                                 void* image, void* bin, uint64_t K_capacity, int* radii_out, float* out,
                                 void* stream)
{
    if (int e = check_dims(P, R, W, H)) return e;
    GsCsi t;
    if (int e = check_csi(f0_hz, df_hz, t)) return e;
    const gs::Dims d = gs::make_dims(P, R, W, H);
    const FwdInputs in{means, scales, rotations, opacities, emission_mlps, viewmatrices, tx_pos, rx_pos,
                       scale_modifier};
    hipStream_t s = (hipStream_t)stream;
    if (int e = forward_geom_impl(d, in, geom, image, radii_out, s)) return e;
    return forward_render_csi_impl(d, t, emission, geom, image, bin, K_capacity, radii_out, out, s);
}

int gsparc_rasterize_backward_csi(int P, int R, int W, int H, double f0_hz, double df_hz, const float* means,
                                  const float* scales, const float* rotations, const float* emission_mlps,
                                  const float* emission, const float* viewmatrices, const float* tx_pos,
                                  const float* rx_pos, float scale_modifier, const int* radii, void* geom,
                                  void* image, void* bin, uint64_t K_capacity, void* grad, const float* dout,
                                  float* dmeans2D, float* dmeans3D, float* demission_mlps, float* dopacity,
                                  float* dcov3D, float* dscales, float* drotations, float* demission, void* stream)
{
    if (int e = check_dims(P, R, W, H)) return e;
    GsCsi t;
    if (int e = check_csi(f0_hz, df_hz, t)) return e;
    const FwdInputs in{means, scales, rotations, nullptr, emission_mlps, viewmatrices, tx_pos, rx_pos,
                       scale_modifier};
    const BwdOutputs out{dmeans2D, dmeans3D, demission_mlps, dopacity, dcov3D, dscales, drotations, demission};
    return backward_impl(GSPARC_GRAD_EXACT, gs::make_dims(P, R, W, H), in, emission, radii, geom, image, bin,
                         K_capacity, grad, dout, nullptr, out, (hipStream_t)stream, &t);
}

int gsparc_autotune(int grad_mode, int P, int R, int W, int H, const float* emission, const int* radii,
                    void* geom, void* image, void* bin, uint64_t K_capacity, void* grad, const float* dout_real,
                    const float* dout_imag, float* out_real, float* out_imag, int iters, int* choice, float* ms,
                    float* table, void* stream)
{
    if (int e = check_dims(P, R, W, H)) return e;
    return autotune_impl(grad_mode, gs::make_dims(P, R, W, H), emission, radii, geom, image, bin, K_capacity, grad,
                         dout_real, dout_imag, out_real, out_imag, iters, choice, ms, table, (hipStream_t)stream);
}

int gsparc_autotune_scenes(int grad_mode, const gsparc_batch* batch, const uint32_t* scene_table, int W, int H,
                           const float* emission, const int* radii, void* geom, void* image, void* bin,
                           uint64_t K_capacity, void* grad, const float* dout_real, const float* dout_imag,
                           float* out_real, float* out_imag, int iters, int* choice, float* ms, float* table,
                           void* stream)
{
    gs::Dims d;
    if (int e = scene_dims(batch, scene_table, W, H, d)) return e;
    return autotune_impl(grad_mode, d, emission, radii, geom, image, bin, K_capacity, grad, dout_real, dout_imag,
                         out_real, out_imag, iters, choice, ms, table, (hipStream_t)stream);
}

int gsparc_autotune_clear(void)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned.clear();
    return GSPARC_OK;
}

size_t gsparc_loss_work_size(int planes, int H, int W)
{
    if (planes <= 0 || H <= 0 || W <= 0) return 0;
    const size_t n = (size_t)planes * H * W;
    return 256 + al(3 * n * sizeof(float)) + 2 * sizeof(float) * (size_t)gs::ssim_blocks(planes, H, W);
}

}  // extern "C"

namespace {

int loss_args(int planes, int H, int W, const float* img, const float* target, float* loss)
{
    if (planes < 0 || H < 0 || W < 0) return fail(GSPARC_ERR_ARG, "negative size");
    if (planes > 65535) return fail(GSPARC_ERR_ARG, "at most 65535 planes per call");
    if ((size_t)planes * H * W == 0) return fail(GSPARC_ERR_ARG, "empty images");
    if (!img || !target || !loss) return fail(GSPARC_ERR_ARG, "null pointer");
    return GSPARC_OK;
}

// The loss's launch sequence on a 256-byte aligned scratch area of at least
// gsparc_loss_work_size - 256 bytes: the three partial-derivative maps, then
// the per-workgroup loss partials.  Every byte it reads is written earlier in
// the same call, so the area needs no clearing.
int loss_l1_ssim_impl(int planes, int H, int W, float lambda_ssim, float C1, float C2, const float* img,
                      const float* target, char* base, float* loss, float* dL_dimg, hipStream_t s)
{
    const size_t n = (size_t)planes * H * W;
    float* dmu1 = dL_dimg ? reinterpret_cast<float*>(base) : nullptr;
    float* ds1 = dL_dimg ? dmu1 + n : nullptr;
    float* ds12 = dL_dimg ? ds1 + n : nullptr;
    float* part = reinterpret_cast<float*>(base + al(3 * n * sizeof(float)));
    const int nb = gs::ssim_blocks(planes, H, W);
    const float l1w = 1.0f - lambda_ssim;
    gs::SsimArgs a{planes, H, W, C1, C2, img, target, nullptr, dmu1, ds1, ds12, part};
    GS_LAUNCH("ssim_fwd", gs::launch_ssim_fwd(a, s), s);
    GS_LAUNCH("loss_reduce", gs::launch_loss_reduce(part, nb, 1.0 / (double)n, l1w, lambda_ssim, loss, s), s);
    if (dL_dimg) {
        // d/dimg of l1w * mean|img - target| + lambda_ssim * (1 - mean SSIM)
        const float inv = (float)(1.0 / (double)n);
        gs::SsimBwdArgs b{planes, H, W, img, target, nullptr, dmu1, ds1, ds12, dL_dimg, -lambda_ssim * inv, l1w * inv};
        GS_LAUNCH("ssim_bwd", gs::launch_ssim_bwd(b, s), s);
    }
    return GSPARC_OK;
}

}  // namespace

// Scratch of gsparc_loss_l1_ssim_noalloc: one array per GPU, created with the
// library's code object on each device and held until the process ends.  In
// host code `scratch` names a host-side stand-in, never the device memory:
// its device address is looked up per device (scratch_ptr) and only that
// address is passed to kernels.
static __device__ __align__(256) unsigned char scratch[GSPARC_LOSS_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// Device address of the current device's `scratch`, looked up on the first
// call on that device (hipGetSymbolAddress with the device current).
int scratch_ptr(char** out)
{
    int dev = -1;
    GS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(GSPARC_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        GS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(GSPARC_ERR_HIP, "hipGetSymbolAddress returned no address for the loss scratch");
        g_scratch_ptr[dev] = p;
    }
    *out = (char*)g_scratch_ptr[dev];
    return GSPARC_OK;
}

}  // namespace

extern "C" {

int gsparc_loss_l1_ssim(int planes, int H, int W, float lambda_ssim, float C1, float C2, const float* img,
                        const float* target, void* work, float* loss, float* dL_dimg, void* stream)
{
    if (int e = loss_args(planes, H, W, img, target, loss)) return e;
    if (!work) return fail(GSPARC_ERR_ARG, "null pointer");
    return loss_l1_ssim_impl(planes, H, W, lambda_ssim, C1, C2, img, target, base_of(work), loss, dL_dimg,
                             (hipStream_t)stream);
}

int gsparc_loss_l1_ssim_noalloc(int planes, int H, int W, float lambda_ssim, float C1, float C2, const float* img,
                                const float* target, float* loss, float* dL_dimg, void* stream)
{
    if (int e = loss_args(planes, H, W, img, target, loss)) return e;
    const size_t need = gsparc_loss_work_size(planes, H, W) - ALIGN;   // the scratch is aligned already
    if (need > GSPARC_LOSS_SCRATCH_BYTES)
        return fail(GSPARC_ERR_ARG, "planes x H x W needs " + std::to_string(need) +
                                        " bytes of loss scratch; the no-allocation loss holds " +
                                        std::to_string(GSPARC_LOSS_SCRATCH_BYTES) +
                                        " (use gsparc_loss_l1_ssim with a workspace)");
    char* base = nullptr;
    if (int e = scratch_ptr(&base)) return e;
    return loss_l1_ssim_impl(planes, H, W, lambda_ssim, C1, C2, img, target, base, loss, dL_dimg,
                             (hipStream_t)stream);
}

size_t gsparc_conf_loss_work_size(int planes, int H, int W)
{
    if (planes <= 0 || H <= 0 || W <= 0) return 0;
    return gsparc_loss_work_size(planes, H, W) + ALIGN + al((size_t)planes * 4 * sizeof(float)) +
           al((size_t)planes * sizeof(float));
}

int gsparc_loss_l1_ssim_conf(int planes, int H, int W, float lambda_ssim, float C1, float C2, float alpha,
                             int hidden, const float* conf_mlp, const float* rx_pos, const float* img,
                             const float* target, void* work, float* loss, float* rec_loss, float* confidence,
                             float* dL_dimg, float* dconf_mlp, void* stream)
{
    if (int e = loss_args(planes, H, W, img, target, loss)) return e;
    if (!work) return fail(GSPARC_ERR_ARG, "null pointer");
    if (conf_mlp && (!rx_pos || hidden < 1 || hidden > 255))
        return fail(GSPARC_ERR_ARG, "the confidence network needs rx_pos and 1..255 hidden units");
    hipStream_t s = (hipStream_t)stream;
    const size_t n = (size_t)planes * H * W;
    char* base = base_of(work);
    float* dmu1 = dL_dimg ? reinterpret_cast<float*>(base) : nullptr;
    float* ds1 = dL_dimg ? dmu1 + n : nullptr;
    float* ds12 = dL_dimg ? ds1 + n : nullptr;
    float* part = reinterpret_cast<float*>(base + al(3 * n * sizeof(float)));
    const int nb = gs::ssim_blocks(planes, H, W);
    char* tail = base + al(al(3 * n * sizeof(float)) + 2 * sizeof(float) * (size_t)nb);
    float* scratch = reinterpret_cast<float*>(tail);
    float* pscale = reinterpret_cast<float*>(tail + al((size_t)planes * 4 * sizeof(float)));
    const float l1w = 1.0f - lambda_ssim;
    gs::SsimArgs a{planes, H, W, C1, C2, img, target, nullptr, dmu1, ds1, ds12, part};
    GS_LAUNCH("ssim_fwd", gs::launch_ssim_fwd(a, s), s);
    gs::ConfLossArgs c;
    c.part = part;
    c.planes = planes;
    c.nbp = gs::ssim_blocks(1, H, W);
    c.hidden = hidden;
    c.inv_np = 1.0 / ((double)H * W);
    c.l1w = l1w;
    c.sw = lambda_ssim;
    c.alpha = alpha;
    c.mlp = conf_mlp;
    c.rx = rx_pos;
    c.loss = loss;
    c.plane_scale = pscale;
    c.rec = rec_loss;
    c.conf = confidence;
    c.dmlp = conf_mlp ? dconf_mlp : nullptr;
    c.scratch = scratch;
    GS_LAUNCH("conf_loss", gs::launch_conf_loss(c, s), s);
    if (dL_dimg) {
        // d/dimg_r of mean_r (C_r L_r - alpha log C_r) = C_r / R dL_r/dimg_r, with
        // dL_r/dimg_r = l1w sign / (H W) - lambda dSSIM_r / (H W)
        const float inv = (float)(1.0 / (double)n);
        gs::SsimBwdArgs b{planes, H, W, img, target, nullptr, dmu1, ds1, ds12, dL_dimg, -lambda_ssim * inv,
                          l1w * inv, pscale};
        GS_LAUNCH("ssim_bwd", gs::launch_ssim_bwd(b, s), s);
    }
    return GSPARC_OK;
}

size_t gsparc_eval_work_size(int planes, int H, int W)
{
    if (planes <= 0 || H <= 0 || W <= 0) return 0;
    return 256 + al(5 * sizeof(float) * (size_t)gs::ssim_blocks(planes, H, W));
}

int gsparc_eval_metrics(int planes, int H, int W, float lambda_ssim, float C1, float C2, int hidden,
                        const float* conf_mlp, const float* rx_pos, const float* img, const float* target, void* work,
                        float* metrics, void* stream)
{
    if (int e = loss_args(planes, H, W, img, target, metrics)) return e;
    if (!work) return fail(GSPARC_ERR_ARG, "null pointer");
    if (conf_mlp && (!rx_pos || hidden < 1 || hidden > 255))
        return fail(GSPARC_ERR_ARG, "the confidence network needs rx_pos and 1..255 hidden units");
    hipStream_t s = (hipStream_t)stream;
    float* part = reinterpret_cast<float*>(base_of(work));
    gs::SsimArgs a{planes, H, W, C1, C2, img, target, nullptr, nullptr, nullptr, nullptr, part};
    a.nsum = 5;
    GS_LAUNCH("ssim_fwd", gs::launch_ssim_fwd(a, s), s);
    gs::EvalArgs e;
    e.part = part;
    e.planes = planes;
    e.nbp = gs::ssim_blocks(1, H, W);
    e.hidden = hidden;
    e.inv_np = 1.0 / ((double)H * W);
    e.lambda_ssim = lambda_ssim;
    e.mlp = conf_mlp;
    e.rx = rx_pos;
    e.metrics = metrics;
    GS_LAUNCH("eval_metrics", gs::launch_eval_metrics(e, s), s);
    return GSPARC_OK;
}

size_t gsparc_csi_loss_work_size(int R, int H, int W)
{
    if (R <= 0 || H <= 0 || W <= 0) return 0;
    const size_t n = (size_t)R * GS_CSI_NC * H * W;
    const size_t npix = (size_t)R * H * W;
    return 256 + 6 * al(n * sizeof(float)) + al(2 * sizeof(float) * (size_t)gs::ssim_blocks(R * GS_CSI_NC, H, W)) +
           al(sizeof(float) * (size_t)gs::csi_loss_blocks(npix)) + gs::loss_reduce2_scratch();
}

int gsparc_csi_loss(int R, int H, int W, float lambda_ssim, float C1, float C2, const float* out, const float* target,
                    void* work, float* loss, float* dL_dout, void* stream)
{
    if (R < 0 || H < 0 || W < 0) return fail(GSPARC_ERR_ARG, "negative size");
    if ((size_t)R * GS_CSI_NC > 65535) return fail(GSPARC_ERR_ARG, "at most 1260 receivers per call");
    const size_t npix = (size_t)R * H * W;
    if (npix == 0) return fail(GSPARC_ERR_ARG, "empty map");
    if (!out || !target || !work || !loss) return fail(GSPARC_ERR_ARG, "null pointer");
    hipStream_t s = (hipStream_t)stream;
    const size_t n = (size_t)R * GS_CSI_NC * H * W;
    char* base = base_of(work);
    float* magA = (float*)base;
    float* magB = (float*)(base + al(n * sizeof(float)));
    float* dmu1 = (float*)(base + 2 * al(n * sizeof(float)));
    float* ds1 = (float*)(base + 3 * al(n * sizeof(float)));
    float* ds12 = (float*)(base + 4 * al(n * sizeof(float)));
    float* dmag = (float*)(base + 5 * al(n * sizeof(float)));
    const int planes = R * GS_CSI_NC;
    const int nbs = gs::ssim_blocks(planes, H, W);
    float* ps = (float*)(base + 6 * al(n * sizeof(float)));
    float* pl = (float*)((char*)ps + al(2 * sizeof(float) * (size_t)nbs));
    const int nbl = gs::csi_loss_blocks(npix);
    double* stage = (double*)((char*)pl + al(sizeof(float) * (size_t)nbl));
    const float l1w = 1.0f - lambda_ssim;
    const double n1 = 2.0 * (double)n;   // real + imaginary parts
    GS_LAUNCH("csi_mag", gs::launch_csi_mag(npix, H * W, out, target, magA, magB, pl, s), s);
    gs::SsimArgs a{planes, H, W, C1, C2, magA, magB, nullptr, dL_dout ? dmu1 : nullptr, dL_dout ? ds1 : nullptr,
                   dL_dout ? ds12 : nullptr, ps};
    GS_LAUNCH("ssim_fwd", gs::launch_ssim_fwd(a, s), s);
    GS_LAUNCH("loss_reduce", gs::launch_loss_reduce2(ps, nbs, 1.0 / (double)n, pl, nbl, 1.0 / n1, l1w, lambda_ssim,
                                                     stage, loss, s), s);
    if (dL_dout) {
        gs::SsimBwdArgs b{planes, H, W, magA, magB, nullptr, dmu1, ds1, ds12, dmag,
                          (float)(-(double)lambda_ssim / (double)n), 0.0f};
        GS_LAUNCH("ssim_bwd", gs::launch_ssim_bwd(b, s), s);
        GS_LAUNCH("csi_grad", gs::launch_csi_grad(npix, H * W, out, target, magA, dmag, (float)((double)l1w / n1),
                                                  dL_dout, s), s);
    }
    return GSPARC_OK;
}

namespace {
int mlp1_args(int P, int R, const float* means, const float* rx_pos, const float* tx_pos, const float* W1,
              const float* b1, const float* W2, const float* b2, const float* w3, const float* b3, void* work,
              gs::Mlp1Args& a)
{
    if (P < 0 || R < 0) return fail(GSPARC_ERR_ARG, "negative size");
    if ((long long)P * R >= (1LL << 31)) return fail(GSPARC_ERR_ARG, "too many (receiver, Gaussian) rows");
    if (P > 0 && (!means || !tx_pos || !W1 || !b1 || !W2 || !b2 || !w3 || !b3 || !work || (R > 0 && !rx_pos)))
        return fail(GSPARC_ERR_ARG, "null pointer");
    a.P = P; a.R = R; a.means = means; a.rx = rx_pos; a.tx = tx_pos; a.w1 = W1; a.b1 = b1; a.w2 = W2; a.b2 = b2;
    a.w3 = w3; a.b3 = b3;
    a.pe1 = reinterpret_cast<float*>(base_of(work));
    a.out = nullptr;
    return GSPARC_OK;
}
}  // namespace

size_t gsparc_mlp1_work_size(int P, int R)
{
    if (P < 0 || R < 0) return 0;
    return 256 + gs::mlp1_work_floats(P, R) * sizeof(float);
}

int gsparc_mlp1_forward(int P, int R, const float* means, const float* rx_pos, const float* tx_pos, const float* W1,
                        const float* b1, const float* W2, const float* b2, const float* w3, const float* b3,
                        void* work, float* emission, void* stream)
{
    gs::Mlp1Args a;
    if (int e = mlp1_args(P, R, means, rx_pos, tx_pos, W1, b1, W2, b2, w3, b3, work, a)) return e;
    if (P == 0) return GSPARC_OK;
    if ((long long)P * R > 0 && !emission) return fail(GSPARC_ERR_ARG, "null pointer");
    a.out = emission;
    GS_LAUNCH("mlp1_fwd", gs::launch_mlp1_fwd(a, (hipStream_t)stream), (hipStream_t)stream);
    return GSPARC_OK;
}

int gsparc_mlp1_backward(int P, int R, const float* means, const float* rx_pos, const float* tx_pos, const float* W1,
                         const float* b1, const float* W2, const float* b2, const float* w3, const float* b3,
                         void* work, const float* demission, float* dmeans, float* grads, void* stream)
{
    gs::Mlp1Args a;
    if (int e = mlp1_args(P, R, means, rx_pos, tx_pos, W1, b1, W2, b2, w3, b3, work, a)) return e;
    if (!grads || (P > 0 && !dmeans) || ((long long)P * R > 0 && !demission))
        return fail(GSPARC_ERR_ARG, "null pointer");
    GS_LAUNCH("mlp1_bwd", gs::launch_mlp1_bwd(a, demission, dmeans, a.pe1, grads, (hipStream_t)stream, false),
              (hipStream_t)stream);
    return GSPARC_OK;
}

int gsparc_mlp1_backward_accumulate(int P, int R, const float* means, const float* rx_pos, const float* tx_pos,
                                    const float* W1, const float* b1, const float* W2, const float* b2,
                                    const float* w3, const float* b3, void* work, const float* demission,
                                    float* dmeans, float* grads, void* stream)
{
    gs::Mlp1Args a;
    if (int e = mlp1_args(P, R, means, rx_pos, tx_pos, W1, b1, W2, b2, w3, b3, work, a)) return e;
    if (!grads || (P > 0 && !dmeans) || ((long long)P * R > 0 && !demission))
        return fail(GSPARC_ERR_ARG, "null pointer");
    GS_LAUNCH("mlp1_bwd", gs::launch_mlp1_bwd(a, demission, dmeans, a.pe1, grads, (hipStream_t)stream, true),
              (hipStream_t)stream);
    return GSPARC_OK;
}

// ---- the training step around the rasterizer
namespace {
bool gaussians_ok(const gsparc_gaussians* g)
{
    return g && g->centre && g->log_scale && g->rotation && g->logit && g->emission_mlps && g->scale &&
           g->rotation_n && g->opacity && g->state && g->state_mlp && g->accum && g->denom && g->max_radius;
}

gs::Gaussians to_dev(const gsparc_gaussians& g)
{
    return gs::Gaussians{g.centre, g.log_scale, g.rotation, g.logit, g.emission_mlps, g.scale, g.rotation_n,
                         g.opacity, g.state, g.state_mlp, g.accum, g.denom, g.max_radius};
}

int check_gaussians(int P, const gsparc_gaussians* g)
{
    if (P < 0) return fail(GSPARC_ERR_ARG, "negative Gaussian count");
    if ((int64_t)P * 388 >= (int64_t)1 << 31) return fail(GSPARC_ERR_ARG, "too many Gaussians");
    if (!gaussians_ok(g)) return fail(GSPARC_ERR_ARG, "null pointer in the Gaussian set");
    return GSPARC_OK;
}
}  // namespace

int gsparc_gaussians_activate(int P, const gsparc_gaussians* g, void* stream)
{
    if (int e = check_gaussians(P, g)) return e;
    GS_LAUNCH("activate", gs::launch_activate(P, to_dev(*g), (hipStream_t)stream), (hipStream_t)stream);
    return GSPARC_OK;
}

int gsparc_adam_step(int P, int groups, const gsparc_adam_config* cfg, const gsparc_gaussians* g,
                     const float* dcentre, const float* dscale, const float* drotation, const float* dopacity,
                     const float* demission_mlps, int n_conf, float* conf, const float* dconf, float* conf_state,
                     void* train_state, void* stream)
{
    if (int e = check_gaussians(P, g)) return e;
    if (!cfg || !train_state) return fail(GSPARC_ERR_ARG, "null pointer");
    if (groups & ~GSPARC_GROUPS_ALL) return fail(GSPARC_ERR_ARG, "unknown parameter group");
    if (((uintptr_t)train_state & 255) != 0) return fail(GSPARC_ERR_ARG, "train_state must be 256-byte aligned");
    if ((groups & GSPARC_GROUP_CENTRE) && !dcentre) return fail(GSPARC_ERR_ARG, "null centre gradient");
    if ((groups & GSPARC_GROUP_SCALING) && !dscale) return fail(GSPARC_ERR_ARG, "null scale gradient");
    if ((groups & GSPARC_GROUP_ROTATION) && !drotation) return fail(GSPARC_ERR_ARG, "null rotation gradient");
    if ((groups & GSPARC_GROUP_OPACITY) && !dopacity) return fail(GSPARC_ERR_ARG, "null opacity gradient");
    if ((groups & GSPARC_GROUP_EMISSION) && !demission_mlps) return fail(GSPARC_ERR_ARG, "null MLP gradient");
    if (n_conf < 0) return fail(GSPARC_ERR_ARG, "negative confidence size");
    if (n_conf > 0 && (groups & GSPARC_GROUP_CONFIDENCE) && (!conf || !dconf || !conf_state))
        return fail(GSPARC_ERR_ARG, "null confidence pointer");
    if (!(cfg->beta1 >= 0.0 && cfg->beta1 < 1.0 && cfg->beta2 >= 0.0 && cfg->beta2 < 1.0 && cfg->eps >= 0.0))
        return fail(GSPARC_ERR_ARG, "Adam betas must lie in [0, 1) and eps be non-negative");
    if (cfg->lr_centre_max_steps <= 0) return fail(GSPARC_ERR_ARG, "lr_centre_max_steps must be positive");
    if ((cfg->lr_centre_init != 0.0 || cfg->lr_centre_final != 0.0) &&
        !(cfg->lr_centre_init > 0.0 && cfg->lr_centre_final > 0.0))
        return fail(GSPARC_ERR_ARG, "the centre schedule needs positive rates (or both zero)");
    gs::AdamCfg c;
    c.beta1 = cfg->beta1;
    c.beta2 = cfg->beta2;
    c.eps = cfg->eps;
    c.lr[gs::AG_CENTRE] = 0.0;
    c.lr[gs::AG_SCALING] = cfg->lr_scaling;
    c.lr[gs::AG_ROTATION] = cfg->lr_rotation;
    c.lr[gs::AG_OPACITY] = cfg->lr_opacity;
    c.lr[gs::AG_EMISSION] = cfg->lr_emission;
    c.lr[gs::AG_CONFIDENCE] = cfg->lr_confidence;
    c.lr_init = cfg->lr_centre_init;
    c.lr_final = cfg->lr_centre_final;
    c.lr_delay_mult = cfg->lr_centre_delay_mult;
    c.lr_delay_steps = cfg->lr_centre_delay_steps;
    c.lr_max_steps = cfg->lr_centre_max_steps;
    c.clamp = cfg->clamp_centres ? 1 : 0;
    for (int k = 0; k < 3; ++k) {
        c.lo[k] = cfg->scene_min[k];
        c.hi[k] = cfg->scene_max[k];
    }
    const gs::GaussianGrads d{dcentre, dscale, drotation, dopacity, demission_mlps};
    uint32_t* counters = (uint32_t*)train_state;
    gs::AdamStep* st = (gs::AdamStep*)((char*)train_state + gs::TRAIN_STEP_OFFSET);
    static_assert(gs::TRAIN_STEP_OFFSET + sizeof(gs::AdamStep) <= GSPARC_TRAIN_STATE_BYTES, "training state size");
    hipStream_t s = (hipStream_t)stream;
    GS_LAUNCH("adam", gs::launch_adam_step(P, groups, c, to_dev(*g), d, (size_t)n_conf, conf, dconf, conf_state,
                                           counters, st, s), s);
    return GSPARC_OK;
}

int gsparc_densify_stats(int P, int R, float grad_scale, const float* dmeans2D, const int* radii,
                         const gsparc_gaussians* g, void* stream)
{
    if (int e = check_gaussians(P, g)) return e;
    if (R < 0 || (int64_t)P * R >= (int64_t)1 << 31) return fail(GSPARC_ERR_ARG, "bad receiver count");
    if ((int64_t)P * R > 0 && (!dmeans2D || !radii)) return fail(GSPARC_ERR_ARG, "null pointer");
    GS_LAUNCH("densify_stats", gs::launch_densify_stats(P, R, grad_scale, dmeans2D, radii, to_dev(*g),
                                                        (hipStream_t)stream), (hipStream_t)stream);
    return GSPARC_OK;
}

int gsparc_opacity_reset(int P, const gsparc_gaussians* g, void* stream)
{
    if (int e = check_gaussians(P, g)) return e;
    GS_LAUNCH("opacity_reset", gs::launch_opacity_reset(P, to_dev(*g), (hipStream_t)stream), (hipStream_t)stream);
    return GSPARC_OK;
}

namespace {
// densify work layout for P Gaussians (gs::DensifyWork)
struct DensifyLayout { size_t flags, keep, order, src_of, block_counts, totals, bytes; };
DensifyLayout densify_layout(int P)
{
    const size_t n2 = 2 * (size_t)P;
    const size_t nb = std::max<size_t>(1, (n2 + 1023) / 1024);
    DensifyLayout L;
    size_t o = 0;
    L.totals = o; o = al(o + 4 * sizeof(uint32_t));
    L.flags = o; o = al(o + (size_t)P);
    L.keep = o; o = al(o + n2);
    L.order = o; o = al(o + n2 * sizeof(uint32_t));
    L.src_of = o; o = al(o + n2 * sizeof(uint32_t));
    L.block_counts = o; o = al(o + 3 * nb * sizeof(uint32_t));
    L.bytes = o + ALIGN;
    return L;
}
}  // namespace

size_t gsparc_densify_work_size(int P)
{
    if (P < 0) return 0;
    return densify_layout(P).bytes;
}

int gsparc_densify_and_prune(int P, const gsparc_densify_config* cfg, const gsparc_gaussians* src,
                             const float* noise, const gsparc_gaussians* dst, int dst_capacity, void* work,
                             int* P_new, int* n_clone, int* n_split, void* stream)
{
    if (int e = check_gaussians(P, src)) return e;
    if (!gaussians_ok(dst) || !cfg || !work || !P_new || !n_clone || !n_split)
        return fail(GSPARC_ERR_ARG, "null pointer");
    if ((int64_t)P * 2 * 388 >= (int64_t)1 << 31) return fail(GSPARC_ERR_ARG, "too many Gaussians");
    if (cfg->densify && P > 0 && !noise) return fail(GSPARC_ERR_ARG, "null noise");
    if (dst_capacity < 0) return fail(GSPARC_ERR_ARG, "negative capacity");
    if (src->centre == dst->centre || src->state == dst->state || src->emission_mlps == dst->emission_mlps)
        return fail(GSPARC_ERR_ARG, "src and dst must be distinct arrays");
    hipStream_t s = (hipStream_t)stream;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    GS_HIP(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone)
        return fail(GSPARC_ERR_ARG, "densify_and_prune reads the new count back: not while capturing");
    *P_new = 0;
    *n_clone = 0;
    *n_split = 0;
    if (P == 0) return GSPARC_OK;
    gs::DensifyCfg c;
    c.densify = cfg->densify ? 1 : 0;
    c.grad_threshold = cfg->grad_threshold;
    c.clone_max_scale = cfg->clone_max_scale;
    c.min_opacity = cfg->min_opacity;
    c.prune_max_scale = cfg->prune_max_scale;
    c.use_bounds = cfg->use_bounds ? 1 : 0;
    for (int k = 0; k < 3; ++k) {
        c.lo[k] = cfg->scene_min[k];
        c.hi[k] = cfg->scene_max[k];
    }
    const DensifyLayout L = densify_layout(P);
    char* base = base_of(work);
    gs::DensifyWork w;
    w.totals = (uint32_t*)(base + L.totals);
    w.flags = (uint8_t*)(base + L.flags);
    w.keep = (uint8_t*)(base + L.keep);
    w.order = (uint32_t*)(base + L.order);
    w.src_of = (uint32_t*)(base + L.src_of);
    w.block_counts = (uint32_t*)(base + L.block_counts);
    const gs::Gaussians p = to_dev(*src), q = to_dev(*dst);
    GS_LAUNCH("densify_lists", gs::launch_densify_lists(P, c, p, noise, w, s), s);
    uint32_t tot[4];
    GS_HIP(hipMemcpyAsync(tot, w.totals, sizeof(tot), hipMemcpyDeviceToHost, s));
    GS_HIP(hipStreamSynchronize(s));
    gs::DensifyPlan pl;
    pl.P = (uint32_t)P;
    pl.n_keep_orig = tot[0];
    pl.n_clone = tot[1];
    pl.n_split = tot[2];
    pl.n_cand = tot[0] + tot[1] + 2 * tot[2];
    pl.P_new = tot[3];
    *P_new = (int)pl.P_new;
    *n_clone = (int)pl.n_clone;
    *n_split = (int)pl.n_split;
    if (pl.n_keep_orig + pl.n_split != (uint32_t)P || pl.P_new > pl.n_cand)
        return fail(GSPARC_ERR_HIP, "densify: inconsistent list lengths");
    if (pl.P_new > (uint32_t)dst_capacity)
        return fail(GSPARC_ERR_CAPACITY, "densify: the new set needs " + std::to_string(pl.P_new) +
                                             " Gaussians, dst holds " + std::to_string(dst_capacity));
    GS_LAUNCH("densify_gather", gs::launch_densify_gather(pl, c, p, q, noise, w, s), s);
    return GSPARC_OK;
}

int gsparc_csi_phase_factors(size_t n, const float* turns, float* cos_out, float* sin_out, void* stream)
{
    if (n > 0 && (!turns || !cos_out || !sin_out)) return fail(GSPARC_ERR_ARG, "null pointer");
    GS_LAUNCH("phase_factors", gs::launch_phase_factors(n, turns, cos_out, sin_out, (hipStream_t)stream),
              (hipStream_t)stream);
    return GSPARC_OK;
}

int gsparc_csi_phasors(size_t n, const float* path_len, double f0_hz, double df_hz, float* out, void* stream)
{
    if (n > 0 && (!path_len || !out)) return fail(GSPARC_ERR_ARG, "null pointer");
    GsCsi t;
    if (int rc = check_csi(f0_hz, df_hz, t)) return rc;
    GS_LAUNCH("csi_phasors", gs::launch_csi_phasors(n, path_len, t, out, (hipStream_t)stream), (hipStream_t)stream);
    return GSPARC_OK;
}

// ---- the spectrum dataset on the device
int gsparc_view_matrices(int N, const double* gateway_quat, const float* rx_pos, float* viewmatrices, void* stream)
{
    if (N < 0) return fail(GSPARC_ERR_ARG, "negative count");
    if (!gateway_quat || (N > 0 && (!rx_pos || !viewmatrices))) return fail(GSPARC_ERR_ARG, "null pointer");
    const double* q = gateway_quat;
    const double n2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    if (!(n2 > 0.0) || !std::isfinite(n2)) return fail(GSPARC_ERR_ARG, "the gateway quaternion must be finite and non-zero");
    GS_LAUNCH("view_matrices", gs::launch_view_matrices(N, q, rx_pos, viewmatrices, (hipStream_t)stream),
              (hipStream_t)stream);
    return GSPARC_OK;
}

int gsparc_gather_batch(int R, int N, int npix, const int* order, int order_len, const void* train_state,
                        const float* spectra, const float* rx_pos, const float* viewmatrices, float* batch_spectra,
                        float* batch_rx, float* batch_views, void* stream)
{
    if (R < 0 || N <= 0 || npix <= 0 || order_len <= 0) return fail(GSPARC_ERR_ARG, "bad size");
    if (R > 65535) return fail(GSPARC_ERR_ARG, "at most 65535 samples per batch");
    if (!order || !train_state || !spectra || !rx_pos || !viewmatrices || !batch_spectra || !batch_rx || !batch_views)
        return fail(GSPARC_ERR_ARG, "null pointer");
    if (((uintptr_t)spectra & 15) || ((uintptr_t)batch_spectra & 15))
        return fail(GSPARC_ERR_ARG, "spectra must be 16-byte aligned");
    GS_LAUNCH("gather_batch", gs::launch_gather_batch(R, N, npix, order, order_len, (const uint32_t*)train_state,
                                                      spectra, rx_pos, viewmatrices, batch_spectra, batch_rx,
                                                      batch_views, (hipStream_t)stream), (hipStream_t)stream);
    return GSPARC_OK;
}

// ---- HIP graphs of the library's own launch sequences
struct gsparc_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int gsparc_graph_begin(void* stream)
{
    if (!stream) return fail(GSPARC_ERR_ARG, "capture needs a stream other than the null stream");
    if (g_prof_on || g_prof_count) return fail(GSPARC_ERR_ARG, "disable profiling before capturing a graph");
    GS_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    return GSPARC_OK;
}

int gsparc_graph_end(void* stream, gsparc_graph** out)
{
    if (!stream || !out) return fail(GSPARC_ERR_ARG, "null pointer");
    *out = nullptr;
    hipGraph_t g = nullptr;
    GS_HIP(hipStreamEndCapture((hipStream_t)stream, &g));
    if (!g) return fail(GSPARC_ERR_HIP, "capture produced no graph");
    hipGraphExec_t e = nullptr;
    const hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    if (err != hipSuccess) {
        hipGraphDestroy(g);
        return fail(GSPARC_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(err));
    }
    gsparc_graph* h = new gsparc_graph;
    h->graph = g;
    h->exec = e;
    *out = h;
    return GSPARC_OK;
}

int gsparc_graph_launch(gsparc_graph* graph, void* stream)
{
    if (!graph || !graph->exec) return fail(GSPARC_ERR_ARG, "null graph");
    GS_HIP(hipGraphLaunch(graph->exec, (hipStream_t)stream));
    return GSPARC_OK;
}

int gsparc_graph_destroy(gsparc_graph* graph)
{
    if (!graph) return GSPARC_OK;
    hipError_t e1 = graph->exec ? hipGraphExecDestroy(graph->exec) : hipSuccess;
    hipError_t e2 = graph->graph ? hipGraphDestroy(graph->graph) : hipSuccess;
    delete graph;
    GS_HIP(e1);
    GS_HIP(e2);
    return GSPARC_OK;
}

int gsparc_profile_enable(int on)
{
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (auto& p : g_prof_pending) {
        g_prof_free.push_back(p.a);
        g_prof_free.push_back(p.b);
    }
    g_prof_pending.clear();
    g_prof_acc.clear();
    if (on < 0 || on > 3) return fail(GSPARC_ERR_ARG, "profile mode must be 0..3");
    const bool count = on == 1 || on == 3;
    if (count && !g_counts_dev) GS_HIP(hipMalloc((void**)&g_counts_dev, 4 * sizeof(unsigned long long)));
    if (count) GS_HIP(hipMemset(g_counts_dev, 0, 4 * sizeof(unsigned long long)));
    g_prof_on = on == 1 || on == 2;
    g_prof_count = count;
    return GSPARC_OK;
}

int gsparc_profile_collect(int* n_kernels)
{
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (auto& p : g_prof_pending) {
        GS_HIP(hipEventSynchronize(p.b));
        float ms = 0.0f;
        GS_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        auto& acc = g_prof_acc[p.name];
        acc.first += ms;
        acc.second += 1;
        g_prof_free.push_back(p.a);
        g_prof_free.push_back(p.b);
    }
    g_prof_pending.clear();
    if (n_kernels) *n_kernels = (int)g_prof_acc.size();
    return GSPARC_OK;
}

int gsparc_profile_query(int i, char* name, int name_len, double* total_ms, long long* launches)
{
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (i < 0 || i >= (int)g_prof_acc.size()) return fail(GSPARC_ERR_ARG, "profile index out of range");
    auto it = g_prof_acc.begin();
    std::advance(it, i);
    if (name && name_len > 0) {
        std::strncpy(name, it->first.c_str(), (size_t)name_len - 1);
        name[name_len - 1] = 0;
    }
    if (total_ms) *total_ms = it->second.first;
    if (launches) *launches = it->second.second;
    return GSPARC_OK;
}

int gsparc_profile_counts(uint64_t* counts)
{
    if (!counts) return fail(GSPARC_ERR_ARG, "null pointer");
    for (int i = 0; i < 4; ++i) counts[i] = 0;
    if (!g_counts_dev) return GSPARC_OK;
    unsigned long long v[4] = {0, 0, 0, 0};
    GS_HIP(hipMemcpy(v, g_counts_dev, sizeof(v), hipMemcpyDeviceToHost));
    for (int i = 0; i < 4; ++i) counts[i] = v[i];
    return GSPARC_OK;
}

}  // extern "C"
