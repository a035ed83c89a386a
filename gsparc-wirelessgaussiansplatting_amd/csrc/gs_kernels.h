// gs_kernels.h -- internal launchers of the gfx950 kernels (C++ linkage; the
// exported C ABI lives in gsparc_abi.cpp and include/gsparc.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gs {

struct Dims {
    int P, R, W, H;
    int gx, gy, T;   // tile grid
    int NB;          // blocks of 256 Gaussians
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

inline Dims make_dims(int P, int R, int W, int H)
{
    Dims d;
    d.P = P; d.R = R; d.W = W; d.H = H;
    d.gx = (W + 15) / 16; d.gy = (H + 15) / 16; d.T = d.gx * d.gy;
    d.NB = (P + 255) / 256;
    return d;
}

struct PreFwdArgs {
    const float* means;      // [P,3]
    const float* scales;     // [P,3]
    const float* rots;       // [P,4]
    const float* opac;       // [P]
    const float* mlp;        // [P*194] or nullptr (emission supplied elsewhere)
    const float* views;      // [R,16] column-major
    const float* tx;         // [3]
    const float* rx;         // [R,3]
    float mod;
    float4* splat;           // [R*P*2]
    float2* sig;             // [R*P] (written only when mlp != nullptr)
    float* depth;            // [R*P]
    int* radii;              // [R*P]
    float* cov3;             // [P*6]
};
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
    uint32_t* skey[2];       // [R*P] depth-sort ping-pong: depth bits
    uint32_t* sval[2];       // [R*P]                       Gaussian index
    uint32_t* shist;         // [R*257*rs_blocks(P)] digit histograms
    uint32_t* ccnt;          // [R*T*nchunks] per-chunk tile counts / bases
    int nchunks, chunk_len;
    uint32_t* plist;         // [Kcap] Gaussian indices in tile-list order
    uint32_t* keys_sorted;   // optional [Kcap] their depth bits
    uint64_t Kcap;
};
int rs_blocks(int P);
int chunk_len_for(const Dims& d);
int chunk_count_for(const Dims& d);
hipError_t launch_depth_sort(const Dims& d, const BinArgs& a, hipStream_t s);
hipError_t launch_chunk_count(const Dims& d, const BinArgs& a, hipStream_t s);
hipError_t launch_chunk_scan(const Dims& d, const BinArgs& a, hipStream_t s);
hipError_t launch_segscan(const Dims& d, const BinArgs& a, hipStream_t s);
hipError_t launch_chunk_scatter(const Dims& d, const BinArgs& a, hipStream_t s);
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
    const float4* splat;     // [R*P*2]
    const float2* sig;       // [R*P]
    float* out_re;           // [R*N]
    float* out_im;           // [R*N]
    float* finalT;           // [R*N]
    uint32_t* ncontrib;      // [R*N]
    const uint32_t* total;   // K as computed on the device
    uint64_t Kcap;           // capacity of plist (nothing is read when K > Kcap)
    unsigned long long* counts;  // optional profiling counters (see below)
};
// Profiling counters (u64 x 4, optional): [0] forward (pixel, Gaussian) pairs
// evaluated, [1] backward pairs evaluated, [2] forward wave-level entry
// tests, [3] backward wave-level entry tests.
// pixels per lane of the blend kernels: 1 (8x8 quarter-tile per wave), 2
// (16x8 half tile) or 4 (16x16 tile per wave); requested = 0 picks by the
// number of tiles
int choose_ppl(const Dims& d, int requested);
hipError_t launch_render_fwd(const Dims& d, const RenderFwdArgs& a, int ppl, hipStream_t s);

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
    float* g8;               // [R*P*8] accumulated: dpx dpy dqa dqb dqc dop dre dim
    int mode;
    const uint32_t* total;
    uint64_t Kcap;
    unsigned long long* counts;  // optional profiling counters
};
hipError_t launch_render_bwd(const Dims& d, const RenderBwdArgs& a, int ppl, hipStream_t s);

struct PreBwdArgs {
    const float* means;
    const float* scales;
    const float* rots;
    const float* mlp;        // or nullptr
    const float* views;
    const float* tx;
    const float* rx;
    float mod;
    const int* radii;        // [R*P]
    const float* g8;         // [R*P*8]
    const float2* sig;       // [R*P] MLP outputs of the forward (used when mlp != nullptr)
    const float* cov3;       // [P*6] cov3D of the forward
    float* dmeans;           // [P,3]
    float* dscales;          // [P,3]
    float* drots;            // [P,4]
    float* dopac;            // [P]
    float* dmlp;             // [P*194] or nullptr
    float* dcov3;            // [P,6] or nullptr
    float* dmeans2d;         // [R*P*3] or nullptr (reference screen-space layout)
    float* dsig;             // [R*P*2] or nullptr (emission gradient)
    int mode;
    float2* dz;              // workspace [R*P]: dL/d(MLP output) per receiver
    float* part;             // workspace [ceil(R/8)*P*12]: per-chunk partial sums
};
hipError_t launch_preprocess_bwd(const Dims& d, const PreBwdArgs& a, hipStream_t s);

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);

struct SsimArgs {
    int planes, H, W;
    float C1, C2;
    const float* img1;
    const float* img2;
    float* map;
    float* dmu1;             // nullptr when not training
    float* ds1;
    float* ds12;
};
hipError_t launch_ssim_fwd(const SsimArgs& a, hipStream_t s);

struct SsimBwdArgs {
    int planes, H, W;
    const float* img1;
    const float* img2;
    const float* dmap;
    const float* dmu1;
    const float* ds1;
    const float* ds12;
    float* dimg1;
};
hipError_t launch_ssim_bwd(const SsimBwdArgs& a, hipStream_t s);

}  // namespace gs
