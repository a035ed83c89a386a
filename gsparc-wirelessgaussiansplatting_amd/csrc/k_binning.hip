// Binning for R receivers: the per-tile lists of Gaussians in (depth, index)
// order.  Together these replace duplicateWithKeys, the 40-bit radix sort of
// (tile, depth) keys and identifyTileRanges (rasterizer_impl.cu:102-171,
// 270-312) and produce the same order: within a tile, Gaussians by ascending
// float depth bits, ties by index (the reference's stable sort sees its
// instances in index order).
//
// Sort first, then bin, so that no per-tile sort is needed:
//   k_os_hist / k_os_pass  (or k_rs_hist / k_rs_scan / k_rs_scatter)
//               stable LSD radix sort of each receiver's visible Gaussians by
//               the 32 depth bits (4 passes of 8-bit digits; input in index
//               order, so equal depths stay in index order): one histogram
//               of all four digits up front, then one launch per pass whose
//               tiles find their offsets by decoupled look-back; for
//               receivers of more than GS_OS_MAX_TILES tiles of 2048, a
//               histogram, a scan and a scatter launch per pass.  The first
//               pass reads depth and radii directly and drops culled
//               Gaussians.
//   k_chunk_count  the sorted sequence of a receiver is cut into chunks; per
//               chunk and tile, the number of instances
//   k_chunk_scan   per (receiver, tile): exclusive prefix over the chunks,
//               total -> segcnt
//   k_segscan   exclusive prefix over (receiver, tile) -> tile ranges, K
//   k_chunk_scatter  one wave per chunk walks it in sorted order, 64
//               Gaussians at a time, and writes each instance at
//               tile base + (instances of the tile before it); the rank
//               among the 64 comes from per-column and per-row ballot masks.
//               The tile lists come out in depth order without a sort.
#include "gs_device.h"
#include "gs_kernels.h"

#include <algorithm>
#include <climits>

namespace gs {

// ---------------------------------------------------------------- scans
// Block-wide exclusive scan of one value per thread (1024 threads max).
template <int NT>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_wave, uint32_t& total)
{
    const int wave = threadIdx.x / GS_WAVE, lane = threadIdx.x % GS_WAVE;
    const uint32_t x = wave_incl_sum(v);
    if (lane == GS_WAVE - 1) s_wave[wave] = x;
    __syncthreads();
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NT / GS_WAVE; ++w) {
        const uint32_t s = s_wave[w];
        if (w < wave) before += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return before + x - v;
}

// K of a workgroup's per-thread partial sums, summed in 64 bits (a thread's
// sum of up to 15 counts below 2^31 can pass 2^32 itself) and saturated:
// 2^32 - 1 stands for "2^32 - 1 or more", which no capacity admits (the host
// keeps capacities below it), so an oversized batch reads as an overflow
// instead of a wrapped count.  Every thread gets the value.
__device__ __forceinline__ uint32_t block_total_saturated(uint64_t v, uint64_t* s_wave64)
{
    const int wave = threadIdx.x / GS_WAVE, lane = threadIdx.x % GS_WAVE;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) s_wave64[wave] = v;
    __syncthreads();
    uint64_t t = 0;
    for (int w = 0; w < (int)(blockDim.x / GS_WAVE); ++w) t += s_wave64[w];
    return t >= 0xffffffffull ? 0xffffffffu : (uint32_t)t;
}

__global__ void __launch_bounds__(1024) k_segscan(int n, const uint32_t* segcnt, uint32_t* segbase, uint2* ranges,
                                                   uint32_t* total)
{
    __shared__ uint32_t s_wave[1024 / GS_WAVE];
    __shared__ uint64_t s_wave64[1024 / GS_WAVE];
    const int per = (n + 1023) / 1024;
    const int i0 = min(n, (int)threadIdx.x * per), i1 = min(n, i0 + per);
    // eight loads at a time (clamped to the thread's range), so that their
    // latencies overlap: one workgroup does the whole scan
    uint32_t sum = 0;
    uint64_t sum64 = 0;
    for (int i = i0; i < i1; i += 8) {
        uint32_t v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = segcnt[min(i + u, i1 - 1)];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            sum += i + u < i1 ? v[u] : 0u;
            sum64 += i + u < i1 ? v[u] : 0u;
        }
    }
    uint32_t tot;
    uint32_t run = block_exclusive_scan<1024>(sum, s_wave, tot);
    tot = block_total_saturated(sum64, s_wave64);
    for (int i = i0; i < i1; i += 8) {
        uint32_t v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = segcnt[min(i + u, i1 - 1)];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (i + u >= i1) break;
            const uint32_t c = v[u];
            segbase[i + u] = run;
            // empty tiles keep the range (0, 0) as after the reference's memset
            // (rasterizer_impl.cu:303)
            ranges[i + u] = c ? make_uint2(run, run + c) : make_uint2(0u, 0u);
            run += c;
        }
    }
    // a depth sort that gave up waiting (total[2], k_os_pass) leaves lists
    // that must not be used: K reads as saturated, i.e. as an overflow
    if (threadIdx.x == 0) total[0] = total[2] ? 0xffffffffu : tot;
}

// The same for n <= GS_SEGSCAN_LDS counts: coalesced loads (all in flight)
// into LDS, the per-thread runs scanned there, coalesced stores.  Run by a
// 1024-thread workgroup.
__device__ __forceinline__ void segscan_lds_body(int n, const uint32_t* segcnt, uint32_t* segbase, uint2* ranges,
                                                 uint32_t* total, uint32_t* s_v, uint32_t* s_wave,
                                                 uint64_t* s_wave64)
{
    constexpr int PER = GS_SEGSCAN_LDS / 1024;
    const int tid = (int)threadIdx.x;
    uint32_t v[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) v[u] = segcnt[min(u * 1024 + tid, n - 1)];
#pragma unroll
    for (int u = 0; u < PER; ++u)
        if (u * 1024 + tid < n) s_v[u * 1024 + tid] = v[u];
    __syncthreads();
    const int per = (n + 1023) / 1024;
    const int i0 = min(n, tid * per), i1 = min(n, i0 + per);
    uint32_t sum = 0;
    uint64_t sum64 = 0;
    for (int i = i0; i < i1; ++i) {
        sum += s_v[i];
        sum64 += s_v[i];
    }
    uint32_t tot;
    uint32_t run = block_exclusive_scan<1024>(sum, s_wave, tot);
    tot = block_total_saturated(sum64, s_wave64);
    for (int i = i0; i < i1; ++i) {
        const uint32_t c = s_v[i];
        s_v[i] = run;
        run += c;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int i = u * 1024 + tid;
        if (i >= n) break;
        const uint32_t b = s_v[i], c = v[u];
        segbase[i] = b;
        // empty tiles keep the range (0, 0) as after the reference's memset
        // (rasterizer_impl.cu:303)
        ranges[i] = c ? make_uint2(b, b + c) : make_uint2(0u, 0u);
    }
    if (tid == 0) total[0] = total[2] ? 0xffffffffu : tot;   // as in segscan_body
}

__global__ void __launch_bounds__(1024) k_segscan_lds(int n, const uint32_t* segcnt, uint32_t* segbase,
                                                       uint2* ranges, uint32_t* total)
{
    __shared__ uint32_t s_v[GS_SEGSCAN_LDS];
    __shared__ uint32_t s_wave[1024 / GS_WAVE];
    __shared__ uint64_t s_wave64[1024 / GS_WAVE];
    segscan_lds_body(n, segcnt, segbase, ranges, total, s_v, s_wave, s_wave64);
}

// ---------------------------------------------------------------- depth sort
// Lanes of the wave holding the same NBITS-bit digit (among valid lanes).
template <int NBITS>
__device__ __forceinline__ uint64_t match_digit(uint32_t dg, bool valid)
{
    uint64_t m = ballot(valid);
#pragma unroll
    for (int b = 0; b < NBITS; ++b) {
        const bool bit = (dg >> b) & 1u;
        const uint64_t v = ballot(bit);
        m &= bit ? v : ~v;
    }
    return m;
}

// Digit of element i of receiver r in pass p.  The first pass reads the
// depth and radius directly: culled Gaussians get the extra digit 256 and
// sort behind every visible one (their positions are never read).
template <bool FIRST>
__device__ __forceinline__ void rs_load(const Dims& d, const BinArgs& a, size_t rP, int i, int pass, uint32_t& key,
                                        uint32_t& val, uint32_t& dg)
{
    if (FIRST) {
        // both loads unconditional (a culled Gaussian's depth is read and
        // discarded), so that a thread's loads can all be in flight at once
        const bool vis = a.radii[rP + i] > 0;
        const uint32_t bits = gs_as_uint(a.depth[rP + i]);
        key = vis ? bits : 0xffffffffu;
        val = (uint32_t)i;
        dg = vis ? (key & 0xffu) : 256u;
    } else {
        key = a.skey[(pass + 1) & 1][rP + i];
        val = a.sval[(pass + 1) & 1][rP + i];
        dg = (key >> (8 * pass)) & 0xffu;
    }
}

// ---- receivers with more than GS_OS_MAX_TILES tiles: three launches per pass
// (per-tile digit histograms, their scan, the scatter), no waits between
// workgroups.  Look-back chains grow with the tile count, and past about 32
// tiles per receiver these measure faster.

template <bool FIRST>
__global__ void __launch_bounds__(GS_RS_THREADS) k_rs_hist(Dims d, BinArgs a, int pass)
{
    constexpr int NB = FIRST ? 257 : 256;
    __shared__ uint32_t h[NB];
    const int r = blockIdx.y, blk = blockIdx.x, nblk = gridDim.x;
    const RxMap rm = rx_map(d, r);
    const int n = FIRST ? (int)rm.P : (int)a.nvis[r];
    const size_t rP = rm.roff;
    for (int i = threadIdx.x; i < NB; i += GS_RS_THREADS) h[i] = 0;
    // the wait-limit flag of the look-back passes, which this path never sets
    if (FIRST && blk == 0 && r == 0 && threadIdx.x == 0) a.total[2] = 0;
    __syncthreads();
    if (blk * GS_RS_TILE < n) {
        // all of the thread's loads first and unconditional (positions past
        // the end read the last element and are masked below), so that their
        // latencies overlap
        uint32_t dg[GS_RS_ITEMS];
#pragma unroll
        for (int k = 0; k < GS_RS_ITEMS; ++k) {
            const int i = min(n - 1, blk * GS_RS_TILE + k * GS_RS_THREADS + (int)threadIdx.x);
            uint32_t key, val;
            rs_load<FIRST>(d, a, rP, i, pass, key, val, dg[k]);
        }
#pragma unroll
        for (int k = 0; k < GS_RS_ITEMS; ++k) {
            const bool valid = blk * GS_RS_TILE + k * GS_RS_THREADS + (int)threadIdx.x < n;
            if (valid) atomicAdd(&h[dg[k]], 1u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < NB; i += GS_RS_THREADS) a.shist[((size_t)r * 257 + i) * nblk + blk] = h[i];
}

// Per receiver: exclusive scan over (digit, block) of the block histograms.
__global__ void __launch_bounds__(1024) k_rs_scan(Dims d, BinArgs a, int nblk, int first)
{
    __shared__ uint32_t s_wave[1024 / GS_WAVE];
    const int r = blockIdx.x;
    const int n = (first ? 257 : 256) * nblk;
    uint32_t* h = a.shist + (size_t)r * 257 * nblk;
    uint32_t carry = 0;
    for (int base = 0; base < n; base += 1024 * 4) {
        uint32_t v[4], s = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int idx = base + threadIdx.x * 4 + j;
            v[j] = idx < n ? h[idx] : 0u;
            s += v[j];
        }
        uint32_t tot;
        uint32_t run = carry + block_exclusive_scan<1024>(s, s_wave, tot);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int idx = base + threadIdx.x * 4 + j;
            if (idx < n) h[idx] = run;
            run += v[j];
        }
        carry += tot;
    }
    __syncthreads();
    // visible Gaussians = everything before the culled digit
    if (first && threadIdx.x == 0) a.nvis[r] = h[(size_t)256 * nblk];
}

// Stable scatter of one pass.  Wave w of the block owns the block's
// elements w*WT .. w*WT+WT-1 (WT = GS_RS_TILE / waves), in GS_RS_ITEMS rounds
// of 64: round ranks come from the digit-match masks, a wave-private running
// count per digit carries them across rounds, and the per-digit prefix over
// the waves joins the waves in input order.  The block then sorts its
// elements into LDS by (digit, input position) and writes each digit's run to
// consecutive global positions, so stores are contiguous per digit rather than
// one element per bucket.  The first pass drops culled Gaussians (digit 256):
// every later pass and the chunk kernels read only the nvis visible ones.
template <bool FIRST>
__global__ void __launch_bounds__(GS_RS_THREADS) k_rs_scatter(Dims d, BinArgs a, int pass)
{
    constexpr int NB = FIRST ? 257 : 256;
    constexpr int NW = GS_RS_THREADS / GS_WAVE;
    constexpr int WT = GS_RS_TILE / NW;
    static_assert(WT == GS_RS_ITEMS * GS_WAVE, "a wave owns GS_RS_ITEMS rounds of 64");
    __shared__ uint32_t s_key[GS_RS_TILE];
    __shared__ uint32_t s_val[GS_RS_TILE];
    __shared__ uint16_t s_dg[GS_RS_TILE];
    __shared__ uint32_t whist[NW][NB];
    __shared__ uint32_t lbase[NB];
    __shared__ uint32_t gbase[NB];
    __shared__ uint32_t s_wave[GS_RS_THREADS / GS_WAVE];
    const int r = blockIdx.y, blk = blockIdx.x, nblk = gridDim.x;
    const RxMap rm = rx_map(d, r);
    const int n = FIRST ? (int)rm.P : (int)a.nvis[r];
    if (blk * GS_RS_TILE >= n) return;
    const size_t rP = rm.roff;
    const int tid = threadIdx.x, wave = tid / GS_WAVE, lane = tid % GS_WAVE;
    for (int i = tid; i < NB; i += GS_RS_THREADS) {
        gbase[i] = a.shist[((size_t)r * 257 + i) * nblk + blk];
#pragma unroll
        for (int w = 0; w < NW; ++w) whist[w][i] = 0;
    }
    __syncthreads();
    uint32_t key[GS_RS_ITEMS], val[GS_RS_ITEMS], dg[GS_RS_ITEMS], rnk[GS_RS_ITEMS];
    const int e0 = wave * WT;
    // all of the thread's loads first and unconditional (positions past the
    // end read the last element and are masked below), so that their
    // latencies overlap
#pragma unroll
    for (int k = 0; k < GS_RS_ITEMS; ++k) {
        const int i = min(n - 1, blk * GS_RS_TILE + e0 + k * GS_WAVE + lane);
        rs_load<FIRST>(d, a, rP, i, pass, key[k], val[k], dg[k]);
    }
#pragma unroll
    for (int k = 0; k < GS_RS_ITEMS; ++k) {
        const bool valid = blk * GS_RS_TILE + e0 + k * GS_WAVE + lane < n;
        const uint64_t m = match_digit<FIRST ? 9 : 8>(dg[k], valid);
        const uint32_t prev = valid ? whist[wave][dg[k]] : 0u;
        rnk[k] = prev + (uint32_t)__popcll(m & lanemask_lt());
        if (valid && (m & lanemask_lt()) == 0) whist[wave][dg[k]] = prev + (uint32_t)__popcll(m);
    }
    __syncthreads();
    // per digit: exclusive prefix over the waves, then over the digits
    uint32_t tot = 0;
    {
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const uint32_t c = whist[w][tid];
            whist[w][tid] = run;
            run += c;
        }
        tot = run;
    }
    uint32_t all;
    const uint32_t lb = block_exclusive_scan<GS_RS_THREADS>(tot, s_wave, all);
    lbase[tid] = lb;
    if (FIRST && tid == 0) {
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const uint32_t c = whist[w][256];
            whist[w][256] = run;
            run += c;
        }
        lbase[256] = all;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < GS_RS_ITEMS; ++k) {
        const int i = blk * GS_RS_TILE + e0 + k * GS_WAVE + lane;
        if (i < n) {
            const uint32_t lp = lbase[dg[k]] + whist[wave][dg[k]] + rnk[k];
            s_key[lp] = key[k];
            s_val[lp] = val[k];
            s_dg[lp] = (uint16_t)dg[k];
        }
    }
    __syncthreads();
    uint32_t* kout = a.skey[pass & 1] + rP;
    uint32_t* vout = a.sval[pass & 1] + rP;
    // the last pass also puts the packed tile rects in depth order, so that
    // the chunk kernels read them in sequence instead of gathering
    const bool rects = !FIRST && pass == 3 && a.rect;
    const int nb = min(GS_RS_TILE, n - blk * GS_RS_TILE);
    for (int e = tid; e < nb; e += GS_RS_THREADS) {
        const uint32_t g = s_dg[e];
        if (FIRST && g == 256) continue;     // culled: never read again
        const uint32_t pos = gbase[g] + (uint32_t)e - lbase[g];
        const uint32_t v = s_val[e];
        kout[pos] = s_key[e];
        vout[pos] = v;
        if (rects) a.srect[rP + pos] = a.rect[rP + v];
    }
}

// ---- receivers with at most GS_OS_MAX_TILES tiles: one launch per pass

// Per receiver, the counts of each of the four 8-bit digits over its visible
// Gaussians (ghist[r][pass][digit]) and the number visible (ghist[r][1024]),
// from the depth and radius in index order: every pass's digit bases are
// known before the first pass runs.  One LDS atomic per (key, digit), then
// one global atomic per nonzero (digit, block) count.
__global__ void __launch_bounds__(GS_RS_THREADS) k_os_hist(Dims d, BinArgs a)
{
    __shared__ uint32_t h[4 * 256];
    __shared__ uint32_t s_vis[GS_RS_THREADS / GS_WAVE];
    const int r = blockIdx.y, blk = blockIdx.x;
    const RxMap rm = rx_map(d, r);
    const int n = (int)rm.P;
    const size_t rP = rm.roff;
    const int tid = threadIdx.x;
    for (int i = tid; i < 4 * 256; i += GS_RS_THREADS) h[i] = 0;
    if (blk == 0 && r == 0 && tid == 0) a.total[2] = 0;   // the passes' wait-limit flag, per sort
    __syncthreads();
    uint32_t nv = 0;
    if (blk * GS_RS_TILE < n) {
        // all of the thread's loads first (clamped positions, masked below)
        uint32_t key[GS_RS_ITEMS], val, dg[GS_RS_ITEMS];
#pragma unroll
        for (int k = 0; k < GS_RS_ITEMS; ++k) {
            const int i = min(n - 1, blk * GS_RS_TILE + k * GS_RS_THREADS + tid);
            rs_load<true>(d, a, rP, i, 0, key[k], val, dg[k]);
        }
#pragma unroll
        for (int k = 0; k < GS_RS_ITEMS; ++k) {
            const bool vis = blk * GS_RS_TILE + k * GS_RS_THREADS + tid < n && dg[k] != 256u;
            if (vis) {
                ++nv;
#pragma unroll
                for (int p = 0; p < 3; ++p) atomicAdd(&h[p * 256 + ((key[k] >> (8 * p)) & 0xffu)], 1u);
            }
            // the top digit (sign and high exponent bits) takes a handful of
            // values over a scene's depths: one atomic per distinct digit of
            // the wave instead of one per key on the same few addresses
            const uint32_t d3 = key[k] >> 24;
            const uint64_t m = match_digit<8>(d3, vis);
            if (vis && (m & lanemask_lt()) == 0) atomicAdd(&h[3 * 256 + d3], (uint32_t)__popcll(m));
        }
    }
    // visible count: wave sums, then one atomic per block
    uint32_t wv = nv;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) wv += __shfl_xor(wv, off);
    if (tid % GS_WAVE == 0) s_vis[tid / GS_WAVE] = wv;
    __syncthreads();
    uint32_t* gh = a.ghist + (size_t)r * GS_OS_GHIST;
    for (int i = tid; i < 4 * 256; i += GS_RS_THREADS)
        if (h[i]) atomicAdd(gh + i, h[i]);
    if (tid == 0) {
        uint32_t t = 0;
#pragma unroll
        for (int w = 0; w < GS_RS_THREADS / GS_WAVE; ++w) t += s_vis[w];
        if (t) atomicAdd(gh + 1024, t);
    }
}

// Status word of a (receiver, tile, digit): published (OS_PUB) by a pass
// (bits 28-29) with the tile's count of the digit, or with OS_INCL the count
// of all tiles up to and including it, in the low 28 bits.  The words are
// zeroed once per sort; a word still holding an earlier pass's tag reads as
// not yet published.
#define OS_PUB (1u << 31)
#define OS_INCL (1u << 30)
#define OS_VAL ((1u << 28) - 1u)
__device__ __forceinline__ uint32_t os_word(int pass, bool incl, uint32_t v)
{
    return OS_PUB | (incl ? OS_INCL : 0u) | ((uint32_t)pass << 28) | v;
}
// a tile waits only on tiles that drew earlier tickets, which are running
// and publish without waiting themselves; the limit bounds the wait anyway
#define OS_SPIN_LIMIT (1 << 20)

// Exclusive prefix, over the tiles before tile blk, of one digit's count:
// walks back through the status words of that digit (st: tile t's word at
// st[t * 256]), OS_LB tiles per round of loads (the tiles of a pass start
// together, so most predecessors show only their own count), adding counts
// until an inclusive prefix.  A round that finds the next word unpublished
// waits and reloads it; past OS_SPIN_LIMIT such rounds it sets *err and
// returns.
#define OS_LB 16
__device__ __forceinline__ uint32_t os_lookback(uint32_t* st, int blk, int pass, uint32_t* err)
{
    uint32_t excl = 0;
    int t = blk - 1;
    int idle = 0;
    const uint32_t tag = OS_PUB | ((uint32_t)pass << 28);
    while (t >= 0) {
        uint32_t w[OS_LB];
#pragma unroll
        for (int k = 0; k < OS_LB; ++k)
            w[k] = t - k >= 0 ? __hip_atomic_load(st + (size_t)(t - k) * 256, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT)
                              : os_word(pass, true, 0u);   // before tile 0: nothing
        int used = 0;
        bool done = false;
#pragma unroll
        for (int k = 0; k < OS_LB; ++k) {
            if (done || used < k) break;
            if ((w[k] & (OS_PUB | (3u << 28))) != tag) break;   // not yet published in this pass
            excl += w[k] & OS_VAL;
            used = k + 1;
            done = (w[k] & OS_INCL) != 0u;
        }
        if (done) break;
        t -= used;
        if (used == 0) {
            if (++idle > OS_SPIN_LIMIT) {
                atomicOr(err, 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    return excl;
}

// One stable pass.  Each workgroup draws a ticket (its tile of the
// receiver's sequence, in launch order) and ranks the tile's keys as
// follows: wave w owns elements w*WT .. w*WT+WT-1 (WT = GS_RS_TILE / waves),
// in GS_RS_ITEMS rounds of 64; round ranks come from the digit-match masks,
// a wave-private running count per digit carries them across rounds, and the
// per-digit prefix over the waves joins the waves in input order.  The
// tile's per-digit counts are published; the counts of the earlier tiles
// come from the look-back, the digit's base in the output from the
// receiver's digit counts (k_os_hist).  The block then sorts its elements
// into LDS by (digit, input position) and writes each digit's run to
// consecutive global positions.  The first pass drops culled Gaussians
// (digit 256): every later pass and the chunk kernels read only the nvis
// visible ones.
template <bool FIRST>
__global__ void __launch_bounds__(GS_RS_THREADS) k_os_pass(Dims d, BinArgs a, int pass, int nblk)
{
    constexpr int NB = FIRST ? 257 : 256;
    constexpr int NW = GS_RS_THREADS / GS_WAVE;
    constexpr int WT = GS_RS_TILE / NW;
    static_assert(WT == GS_RS_ITEMS * GS_WAVE, "a wave owns GS_RS_ITEMS rounds of 64");
    static_assert(GS_RS_THREADS == 256, "one thread per digit");
    __shared__ uint32_t s_key[GS_RS_TILE];
    __shared__ uint32_t s_val[GS_RS_TILE];
    __shared__ uint16_t s_dg[GS_RS_TILE];
    __shared__ uint32_t whist[NW][NB];
    __shared__ uint32_t lbase[NB];
    __shared__ uint32_t gbase[NB];
    __shared__ uint32_t s_wave[GS_RS_THREADS / GS_WAVE];
    __shared__ int s_tile;
    const int r = blockIdx.y;
    const RxMap rm = rx_map(d, r);
    const uint32_t* gh = a.ghist + (size_t)r * GS_OS_GHIST;
    const int n = FIRST ? (int)rm.P : (int)a.nvis[r];
    const int tid = threadIdx.x, wave = tid / GS_WAVE, lane = tid % GS_WAVE;
    if (tid == 0) s_tile = (int)atomicAdd(a.tick + (size_t)pass * d.R + r, 1u);
    for (int i = tid; i < NB; i += GS_RS_THREADS) {
#pragma unroll
        for (int w = 0; w < NW; ++w) whist[w][i] = 0;
    }
    __syncthreads();
    const int blk = s_tile;
    // every receiver's visible count is written, also where the pass has no
    // tile (a scene without Gaussians): the later passes and the chunk
    // kernels take it as their element count
    if (FIRST && blk == 0 && tid == 0) a.nvis[r] = gh[1024];
    if (blk * GS_RS_TILE >= n) return;   // uniform over the block
    const size_t rP = rm.roff;
    uint32_t key[GS_RS_ITEMS], val[GS_RS_ITEMS], dg[GS_RS_ITEMS], rnk[GS_RS_ITEMS];
    const int e0 = wave * WT;
    // all of the thread's loads first and unconditional (positions past the
    // end read the last element and are masked below), so that their
    // latencies overlap
#pragma unroll
    for (int k = 0; k < GS_RS_ITEMS; ++k) {
        const int i = min(n - 1, blk * GS_RS_TILE + e0 + k * GS_WAVE + lane);
        rs_load<FIRST>(d, a, rP, i, pass, key[k], val[k], dg[k]);
    }
    // the receiver's count of this thread's digit in this pass (read early)
    const uint32_t gcount = gh[pass * 256 + tid];
#pragma unroll
    for (int k = 0; k < GS_RS_ITEMS; ++k) {
        const bool valid = blk * GS_RS_TILE + e0 + k * GS_WAVE + lane < n;
        const uint64_t m = match_digit<FIRST ? 9 : 8>(dg[k], valid);
        const uint32_t prev = valid ? whist[wave][dg[k]] : 0u;
        rnk[k] = prev + (uint32_t)__popcll(m & lanemask_lt());
        if (valid && (m & lanemask_lt()) == 0) whist[wave][dg[k]] = prev + (uint32_t)__popcll(m);
    }
    __syncthreads();
    // per digit: exclusive prefix over the waves, then over the digits
    uint32_t tot = 0;
    {
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const uint32_t c = whist[w][tid];
            whist[w][tid] = run;
            run += c;
        }
        tot = run;
    }
    // publish this tile's count of digit tid (tile 0: already inclusive)
    uint32_t* st = a.lbstat + (size_t)r * nblk * 256 + tid;
    __hip_atomic_store(st + (size_t)blk * 256, os_word(pass, blk == 0, tot), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    uint32_t all;
    const uint32_t lb = block_exclusive_scan<GS_RS_THREADS>(tot, s_wave, all);
    lbase[tid] = lb;
    if (FIRST && tid == 0) {
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const uint32_t c = whist[w][256];
            whist[w][256] = run;
            run += c;
        }
        lbase[256] = all;
    }
    // the digit's first output position: digits below it over the whole
    // receiver, then the digit's elements in earlier tiles
    uint32_t gall;
    const uint32_t db = block_exclusive_scan<GS_RS_THREADS>(gcount, s_wave, gall);
    uint32_t excl = 0;
    if (blk > 0) {
        excl = os_lookback(st, blk, pass, a.total + 2);
        __hip_atomic_store(st + (size_t)blk * 256, os_word(pass, true, excl + tot), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    gbase[tid] = db + excl;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < GS_RS_ITEMS; ++k) {
        const int i = blk * GS_RS_TILE + e0 + k * GS_WAVE + lane;
        if (i < n) {
            const uint32_t lp = lbase[dg[k]] + whist[wave][dg[k]] + rnk[k];
            s_key[lp] = key[k];
            s_val[lp] = val[k];
            s_dg[lp] = (uint16_t)dg[k];
        }
    }
    __syncthreads();
    uint32_t* kout = a.skey[pass & 1] + rP;
    uint32_t* vout = a.sval[pass & 1] + rP;
    // the last pass also puts the packed tile rects in depth order, so that
    // the chunk kernels read them in sequence instead of gathering
    const bool rects = !FIRST && pass == 3 && a.rect;
    const int nb = min(GS_RS_TILE, n - blk * GS_RS_TILE);
    for (int e = tid; e < nb; e += GS_RS_THREADS) {
        const uint32_t g = s_dg[e];
        if (FIRST && g == 256) continue;     // culled: never read again
        const uint32_t pos = gbase[g] + (uint32_t)e - lbase[g];
        const uint32_t v = s_val[e];
        kout[pos] = s_key[e];
        vout[pos] = v;
        if (rects) a.srect[rP + pos] = a.rect[rP + v];
    }
}

// ---------------------------------------------------------------- chunks
// Tile rect of sorted position i (Gaussian g) of the receiver at rP.
__device__ __forceinline__ void chunk_rect(const Dims& d, const BinArgs& a, size_t rP, int i, uint32_t g, int& x0,
                                           int& y0, int& x1, int& y1)
{
    if (a.rect) {
        const uint32_t p = a.srect[rP + i];
        x0 = (int)(p & 0xffu);
        y0 = (int)((p >> 8) & 0xffu);
        x1 = (int)((p >> 16) & 0xffu);
        y1 = (int)(p >> 24);
        return;
    }
    const int radius = a.radii[rP + g];
    x0 = y0 = x1 = y1 = 0;
    if (radius > 0) {
        const float4 s0 = a.splat[2 * (rP + g)];
        get_rect(s0.x, s0.y, radius, d.gx, d.gy, x0, y0, x1, y1);
    }
}

// per-tile counters of a chunk in LDS (else in the global count matrix); the
// bound keeps the GS_SCATTER_WAVES slices of a scatter workgroup within LDS
__host__ __device__ inline bool chunk_lds(const Dims& d)
{
    return d.T <= GS_CHUNK_LDS_TILES && (size_t)(d.gx + d.gy) * 8 + (size_t)d.T * 4 <= 122880 / GS_SCATTER_WAVES;
}

// u64 words of one wave's dynamic LDS in k_chunk_scatter
__host__ __device__ inline size_t chunk_scatter_slice(const Dims& d)
{
    return (size_t)(d.gx + d.gy) + (chunk_lds(d) ? ((size_t)d.T + 1) / 2 : 0);
}

// Orders the LDS accesses of one wave (its lanes exchange data through LDS
// and no other wave touches its slice).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// instances per (chunk, tile) -> ccnt[(r*T + t) * nchunks + c]
template <int CPB>
__global__ void __launch_bounds__(GS_RS_THREADS) k_chunk_count(Dims d, BinArgs a)
{
    // LDS path: a difference array per tile row, (gx + 1) wide, per chunk: a
    // rect adds +1 at (y, x0) and -1 at (y, x1) for each of its rows; each
    // tile's count is then the prefix of its row up to its column.  The
    // workgroup takes CPB consecutive chunks (fewer, fuller workgroups).
    extern __shared__ uint32_t s_cnt[];
    const int cb = blockIdx.x * CPB, r = blockIdx.y;
    const bool lds = chunk_lds(d);
    const int n = (int)a.nvis[r];
    const size_t rP = rx_map(d, r).roff;
    const int rw = d.gx + 1, ra = rw * d.gy;   // one chunk's difference arrays
    // the receiver's own chunks only: the counts of chunks past its visible
    // Gaussians are never read (the scan and the placement stop there too)
    const int nch = (n + a.chunk_len - 1) / a.chunk_len;
    if (cb >= nch) return;   // uniform over the workgroup
    const int nc = min(CPB, nch - cb);         // chunks of this workgroup (>= 1)
    uint32_t* slot = a.ccnt + (size_t)r * d.T * a.nchunks + cb;   // chunk cb + q, tile t at slot[t * nchunks + q]
    if (lds) {
        for (int t = threadIdx.x; t < nc * ra; t += GS_RS_THREADS) s_cnt[t] = 0;
    } else {
        for (int t = threadIdx.x; t < nc * d.T; t += GS_RS_THREADS)
            slot[(size_t)(t % d.T) * a.nchunks + t / d.T] = 0;
    }
    __syncthreads();
    for (int q = 0; q < nc; ++q) {
        const int i0 = (cb + q) * a.chunk_len, i1 = min(n, i0 + a.chunk_len);
        uint32_t* sq = s_cnt + q * ra;
        for (int i = i0 + threadIdx.x; i < i1; i += GS_RS_THREADS) {
            int x0, y0, x1, y1;
            chunk_rect(d, a, rP, i, a.rect ? 0u : a.sval[1][rP + i], x0, y0, x1, y1);
            if (x0 >= x1) continue;
            for (int ty = y0; ty < y1; ++ty) {
                if (lds) {
                    atomicAdd(&sq[ty * rw + x0], 1u);
                    atomicAdd(&sq[ty * rw + x1], 0xffffffffu);   // -1 modulo 2^32
                } else {
                    for (int tx = x0; tx < x1; ++tx) atomicAdd(&slot[(size_t)(ty * d.gx + tx) * a.nchunks + q], 1u);
                }
            }
        }
    }
    if (lds) {
        __syncthreads();
        for (int t = threadIdx.x; t < nc * d.T; t += GS_RS_THREADS) {
            const int q = t / d.T, tt = t % d.T;
            const int ty = tt / d.gx, tx = tt % d.gx;
            const uint32_t* sq = s_cnt + q * ra + ty * rw;
            uint32_t v = 0;
            for (int x = 0; x <= tx; ++x) v += sq[x];
            slot[(size_t)tt * a.nchunks + q] = v;
        }
    }
}

// per (receiver, tile): exclusive prefix over chunks, one wave per row; the
// total is the tile's instance count
__global__ void __launch_bounds__(256) k_chunk_scan(Dims d, BinArgs a)
{
    const int row = blockIdx.x * (256 / GS_WAVE) + threadIdx.x / GS_WAVE;
    const int lane = threadIdx.x % GS_WAVE;
    if (row >= d.R * d.T) return;
    uint32_t* cnt = a.ccnt + (size_t)row * a.nchunks;
    // the receiver's chunks (k_chunk_count writes no others)
    const int nch = ((int)a.nvis[row / d.T] + a.chunk_len - 1) / a.chunk_len;
    uint32_t carry = 0;
    // four 64-chunk steps loaded at a time, so that their latencies overlap
    for (int cb = 0; cb < nch; cb += 4 * GS_WAVE) {
        uint32_t v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = cnt[min(cb + u * GS_WAVE + lane, nch - 1)];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int c = cb + u * GS_WAVE + lane;
            if (cb + u * GS_WAVE >= nch) break;
            const uint32_t vc = c < nch ? v[u] : 0u;
            const uint32_t x = wave_incl_sum(vc);
            if (c < nch) cnt[c] = carry + x - vc;
            carry += __builtin_amdgcn_readlane(x, GS_WAVE - 1);
        }
    }
    if (lane == 0) a.segcnt[row] = carry;
}

// One wave per chunk: order-preserving placement of the chunk's instances.
// The chunk is walked 64 sorted Gaussians (lanes) at a time; the batch's
// (Gaussian, tile) pairs are spread evenly over the lanes, so one large
// Gaussian does not serialise its batch.  An instance of lane i in tile t goes
// to base(t) + (lanes below i covering t); the lanes covering t are the AND of
// the column and row ballot masks.
// A workgroup holds GS_SCATTER_WAVES independent waves (chunks
// blockIdx.x * GS_SCATTER_WAVES + wave), each with its own LDS slice and no
// workgroup barrier, so that more waves are resident per CU.
// Dynamic LDS per wave: column masks [gx], row masks [gy] (u64), then (when
// the tiles fit) the running base of every tile [T] (u32).
__global__ void __launch_bounds__(GS_WAVE * GS_SCATTER_WAVES) k_chunk_scatter(Dims d, BinArgs a)
{
    extern __shared__ uint64_t s_dyn[];
    __shared__ uint32_t s_end_w[GS_SCATTER_WAVES][GS_WAVE];   // inclusive prefix of the pair counts
    __shared__ int4 s_own_w[GS_SCATTER_WAVES][GS_WAVE];       // x0, y0, width, Gaussian index
    __shared__ uint32_t s_key_w[GS_SCATTER_WAVES][GS_WAVE];
    __shared__ int s_mark_w[GS_SCATTER_WAVES][GS_WAVE];       // lane whose pairs start at this window position
    __shared__ float s_inv_w[GS_SCATTER_WAVES][GS_WAVE];      // 1 / rect width
    const int wave = (int)threadIdx.x / GS_WAVE;
    uint32_t* s_end = s_end_w[wave];
    int4* s_own = s_own_w[wave];
    uint32_t* s_key = s_key_w[wave];
    int* s_mark = s_mark_w[wave];
    float* s_inv = s_inv_w[wave];
    uint64_t* s_mask = s_dyn + (size_t)wave * chunk_scatter_slice(d);
    uint64_t* s_mx = s_mask;
    uint64_t* s_my = s_mask + d.gx;
    uint32_t* s_base = (uint32_t*)(s_mask + d.gx + d.gy);
    if ((uint64_t)a.total[0] > a.Kcap) return;   // capacity overflow: reported by the host
    const int c = (int)blockIdx.x * GS_SCATTER_WAVES + wave, r = blockIdx.y;
    if (c >= a.nchunks) return;
    const int n = (int)a.nvis[r];
    const int i0 = c * a.chunk_len;
    if (i0 >= n) return;
    const int i1 = min(n, i0 + a.chunk_len);
    const bool lds = chunk_lds(d);
    const int lane = (int)threadIdx.x % GS_WAVE;
    const size_t rP = rx_map(d, r).roff;
    const size_t rT = (size_t)r * d.T;
    uint32_t* slot = a.ccnt + rT * a.nchunks + c;
    for (int t = lane; t < d.T; t += GS_WAVE) {
        const uint32_t b = a.segbase[rT + t] + slot[(size_t)t * a.nchunks];
        if (lds) s_base[t] = b;
        else slot[(size_t)t * a.nchunks] = b;
    }
    for (int b0 = i0; b0 < i1; b0 += GS_WAVE) {
        const int i = b0 + lane;
        int x0 = 0, y0 = 0, x1 = 0, y1 = 0;
        uint32_t g = 0, dbits = 0;
        if (i < i1) {
            g = a.sval[1][rP + i];
            dbits = a.skey[1][rP + i];
            chunk_rect(d, a, rP, i, g, x0, y0, x1, y1);
        }
        const bool ne = x0 < x1 && y0 < y1;
        const uint32_t area = ne ? (uint32_t)((x1 - x0) * (y1 - y0)) : 0u;
        const uint32_t incl = wave_incl_sum(area);
        const uint32_t npairs = __builtin_amdgcn_readlane(incl, GS_WAVE - 1);
        int ux0 = ne ? x0 : INT_MAX, uy0 = ne ? y0 : INT_MAX, ux1 = ne ? x1 : 0, uy1 = ne ? y1 : 0;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            ux0 = min(ux0, __shfl_xor(ux0, off));
            uy0 = min(uy0, __shfl_xor(uy0, off));
            ux1 = max(ux1, __shfl_xor(ux1, off));
            uy1 = max(uy1, __shfl_xor(uy1, off));
        }
        ux0 = __builtin_amdgcn_readfirstlane(ux0);
        uy0 = __builtin_amdgcn_readfirstlane(uy0);
        ux1 = __builtin_amdgcn_readfirstlane(ux1);
        uy1 = __builtin_amdgcn_readfirstlane(uy1);
        wave_sync();   // the previous batch is done with the LDS tables
        // lanes (Gaussians of this batch) covering each column / row
        for (int tx = ux0; tx < ux1; ++tx) {
            const uint64_t m = ballot(ne && x0 <= tx && tx < x1);
            if (lane == 0) s_mx[tx] = m;
        }
        for (int ty = uy0; ty < uy1; ++ty) {
            const uint64_t m = ballot(ne && y0 <= ty && ty < y1);
            if (lane == 0) s_my[ty] = m;
        }
        s_end[lane] = incl;
        s_own[lane] = make_int4(x0, y0, x1 - x0, (int)g);
        s_inv[lane] = ne ? 1.0f / (float)(x1 - x0) : 0.0f;
        s_key[lane] = dbits;
        wave_sync();
        // place the instances, 64 pairs per window: pair p of the batch
        // belongs to the last lane with a nonempty rect starting at or before
        // p (start markers in LDS, a wave max-scan, the owner carried across
        // windows); its rank inside the tile is the number of lower lanes of
        // the batch covering the same tile
        const uint32_t first = incl - area;
        int carry = 0;
        for (uint32_t base = 0; base < npairs; base += GS_WAVE) {
            s_mark[lane] = -1;
            __builtin_amdgcn_wave_barrier();
            if (area > 0 && first >= base && first < base + GS_WAVE) s_mark[first - base] = lane;
            __builtin_amdgcn_wave_barrier();
            const int lo = max(carry, wave_incl_max(s_mark[lane]));
            carry = __builtin_amdgcn_readlane(lo, GS_WAVE - 1);
            __builtin_amdgcn_wave_barrier();
            const uint32_t p = base + lane;
            if (p >= npairs) continue;
            const int4 ow = s_own[lo];
            const uint32_t q = p - (lo ? s_end[lo - 1] : 0u);   // pair index within the owner's rect
            // row = floor(q / w) as floor((q + 1/2) / w) in float: the half
            // keeps the quotient 1/(2w) away from every integer, more than the
            // rounding error while q + w < 2^20 (the host checks T < 2^20)
            const uint32_t qy = (uint32_t)fmaf((float)q, s_inv[lo], 0.5f * s_inv[lo]);
            const int ty = ow.y + (int)qy;
            const int tx = ow.x + (int)(q - qy * (uint32_t)ow.z);
            const int t = ty * d.gx + tx;
            const uint64_t m = s_mx[tx] & s_my[ty];
            const uint32_t b = lds ? s_base[t] : slot[(size_t)t * a.nchunks];
            const uint32_t pos = b + (uint32_t)__popcll(m & ((1ull << lo) - 1ull));
            a.plist[pos] = (uint32_t)ow.w;
            if (a.keys_sorted) a.keys_sorted[pos] = s_key[lo];
        }
        wave_sync();
        // advance the base of every tile of the batch's union rect by the
        // number of the batch's Gaussians covering it (one lane per tile)
        // (an empty batch has ux0 = uy0 = INT_MAX: no tiles)
        const int uw = npairs ? ux1 - ux0 : 0, uh = npairs ? uy1 - uy0 : 0;
        for (int q = lane; q < uw * uh; q += GS_WAVE) {
            const int ty = uy0 + q / uw, tx = ux0 + q % uw;
            const uint32_t cnt = (uint32_t)__popcll(s_mx[tx] & s_my[ty]);
            if (cnt == 0) continue;
            const int t = ty * d.gx + tx;
            if (lds) s_base[t] += cnt;
            else slot[(size_t)t * a.nchunks] += cnt;
        }
        wave_sync();
    }
}

// ---------------------------------------------------------------- row placement
// For maps of up to GS_ROW_PLACE_MAX_T tiles the placement runs one wave per
// (chunk, tile row).  The wave loads the chunk's depth-ordered packed rects
// (NI groups of 64) and compacts the Gaussians that reach its row, in order,
// through LDS; then, per tile of the row, the members covering it
// (x0 <= tx < x1) are one ballot per 64-member group, their ranks popcounts
// of it, and each goes to the tile's base for the chunk + its rank: the lists
// of k_chunk_scatter, with each store writing consecutive positions.  Tiles
// outside the union of the members' rects are skipped.
template <int NI, bool KEYS>
__global__ void __launch_bounds__(GS_WAVE * GS_RP_WAVES) k_row_place(Dims d, BinArgs a)
{
    __shared__ uint32_t s_rc[GS_RP_WAVES][NI * GS_WAVE];
    __shared__ uint32_t s_gv[GS_RP_WAVES][NI * GS_WAVE];
    __shared__ uint32_t s_kv[GS_RP_WAVES][KEYS ? NI * GS_WAVE : 1];
    if ((uint64_t)a.total[0] > a.Kcap) return;   // capacity overflow: reported by the host
    const int r = blockIdx.y;
    const int w = (int)threadIdx.x / GS_WAVE;
    const int idx = (int)blockIdx.x * GS_RP_WAVES + w;
    const int c = idx / d.gy, ty = idx % d.gy;
    if (c >= a.nchunks) return;
    const int lane = (int)threadIdx.x % GS_WAVE;
    const int n = (int)a.nvis[r];
    const int i0 = c * a.chunk_len;
    const int m = min(n - i0, NI * GS_WAVE);   // Gaussians of this chunk
    if (m <= 0) return;
    const size_t rP = rx_map(d, r).roff;
    const size_t rowT = (size_t)r * d.T + (size_t)ty * d.gx;
    // every load first (positions past the end read the last Gaussian and
    // are masked below), so that their latencies overlap
    uint32_t rc[NI], gv[NI], kv[NI];
#pragma unroll
    for (int k = 0; k < NI; ++k) {
        const size_t i = rP + (size_t)(i0 + min(k * GS_WAVE + lane, m - 1));
        rc[k] = a.srect[i];
        gv[k] = a.sval[1][i];
        kv[k] = KEYS ? a.skey[1][i] : 0u;
    }
    // members: Gaussians of the chunk whose rect reaches row ty, in order
    int nm = 0;
#pragma unroll
    for (int k = 0; k < NI; ++k) {
        const uint32_t p = rc[k];
        const int y0 = (int)((p >> 8) & 0xffu), y1 = (int)(p >> 24);
        const bool in = k * GS_WAVE + lane < m && y0 <= ty && ty < y1 && (p & 0xffu) < ((p >> 16) & 0xffu);
        const uint64_t mk = ballot(in);
        if (in) {
            const int j = nm + (int)__popcll(mk & lanemask_lt());
            s_rc[w][j] = p;
            s_gv[w][j] = gv[k];
            if (KEYS) s_kv[w][j] = kv[k];
        }
        nm += (int)__popcll(mk);
    }
    if (nm == 0) return;
    wave_sync();
    int x0[NI], x1[NI];
    // union of the members' columns: the largest 255 - x0 and x1 (-1 for a
    // lane without members, the identity of wave_incl_max)
    int lo = -1, hi = -1;
#pragma unroll
    for (int g = 0; g < NI; ++g) {
        x0[g] = x1[g] = 0;   // an empty column range past the members
        if (g * GS_WAVE >= nm) continue;
        const int j = g * GS_WAVE + lane;
        if (j < nm) {
            const uint32_t p = s_rc[w][j];
            x0[g] = (int)(p & 0xffu);
            x1[g] = (int)((p >> 16) & 0xffu);
            gv[g] = s_gv[w][j];
            if (KEYS) kv[g] = s_kv[w][j];
            lo = max(lo, 255 - x0[g]);
            hi = max(hi, x1[g]);
        }
    }
    const int ux0 = 255 - __builtin_amdgcn_readlane(wave_incl_max(lo), GS_WAVE - 1);
    const int ux1 = __builtin_amdgcn_readlane(wave_incl_max(hi), GS_WAVE - 1);
    for (int tb = ux0; tb < ux1; tb += GS_WAVE) {
        const int tl = tb + lane;
        const uint32_t bl = tl < ux1 ? a.segbase[rowT + tl] + a.ccnt[(rowT + tl) * a.nchunks + c] : 0u;
        const int te = min(tb + GS_WAVE, ux1);
        for (int tx = tb; tx < te; ++tx) {
            uint32_t base = (uint32_t)__builtin_amdgcn_readlane((int)bl, tx - tb);
#pragma unroll
            for (int g = 0; g < NI; ++g) {
                if (g * GS_WAVE >= nm) break;
                const bool cv = x0[g] <= tx && tx < x1[g];
                const uint64_t mk = ballot(cv);
                if (cv) {
                    const uint32_t pos = base + (uint32_t)__popcll(mk & lanemask_lt());
                    a.plist[pos] = gv[g];
                    if (KEYS) a.keys_sorted[pos] = kv[g];
                }
                base += (uint32_t)__popcll(mk);
            }
        }
    }
}

// ---------------------------------------------------------------- tile order
// 128 buckets of list length: 4 per octave (leading bit + the next two).
__device__ __forceinline__ int length_bucket(uint32_t L)
{
    if (L < 4) return (int)L;
    const int e = 31 - __clz(L);
    return 4 * e + (int)((L >> (e - 2)) & 3u);
}

// One workgroup per list l: the tiles of receivers l, l + L, l + 2L, ...
// (L = number of lists) by descending length bucket, into order + l * stride;
// the rest of the list's slots (up to stride, or up to len for a single list)
// get the empty marker 0xffffffff.  Workgroup 0 also records the mode in
// total[1].
// List l of L (a 1024-thread workgroup); len_of(rt) is the list length of
// (receiver, tile) rt.
template <class LenOf>
__device__ __forceinline__ void tile_order_body(LenOf len_of, int R, int T, int stride, uint32_t* order,
                                                uint32_t* total, int mode, int len, int l, int L, uint32_t* cnt)
{
    const int n = (R - l + L - 1) / L * T;
    uint32_t* out = order + (size_t)l * stride;
    if (l == 0 && threadIdx.x == 0) total[1] = (uint32_t)mode;
    const int pad_end = stride ? stride : len;
    for (int i = n + (int)threadIdx.x; i < pad_end; i += 1024) out[i] = 0xffffffffu;
    for (int i = threadIdx.x; i < 128; i += 1024) cnt[i] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 1024)
        atomicAdd(&cnt[length_bucket(len_of((size_t)(l + (i / T) * L) * T + i % T))], 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int b = 127; b >= 0; --b) {
            const uint32_t c = cnt[b];
            cnt[b] = run;
            run += c;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 1024) {
        const uint32_t rt = (uint32_t)((l + (i / T) * L) * T + i % T);
        out[atomicAdd(&cnt[length_bucket(len_of(rt))], 1u)] = rt;
    }
}

__global__ void __launch_bounds__(1024) k_tile_order(const uint2* ranges, int R, int T, int stride, uint32_t* order,
                                                     uint32_t* total, int mode, int len)
{
    __shared__ uint32_t cnt[128];
    tile_order_body([ranges](size_t rt) { const uint2 rg = ranges[rt]; return rg.y - rg.x; }, R, T, stride, order,
                    total, mode, len, blockIdx.x, gridDim.x, cnt);
}

// Workgroup 0 scans the (receiver, tile) counts (k_segscan_lds); workgroups
// 1 .. L build the L work-order lists from the same counts (a list's length
// is its count), so both run in one launch.
__global__ void __launch_bounds__(1024) k_segscan_order(int n, const uint32_t* segcnt, uint32_t* segbase,
                                                         uint2* ranges, uint32_t* total, int R, int T, int stride,
                                                         uint32_t* order, int mode, int len)
{
    __shared__ uint32_t s_v[GS_SEGSCAN_LDS];
    __shared__ uint32_t s_wave[1024 / GS_WAVE];
    __shared__ uint64_t s_wave64[1024 / GS_WAVE];
    if (blockIdx.x == 0) {
        segscan_lds_body(n, segcnt, segbase, ranges, total, s_v, s_wave, s_wave64);
        return;
    }
    tile_order_body([segcnt](size_t rt) { return segcnt[rt]; }, R, T, stride, order, total, mode, len,
                    (int)blockIdx.x - 1, (int)gridDim.x - 1, s_v);
}

// Mode 3: one workgroup per receiver r sorts its own tiles into slot r / 8
// of XCD r % 8; workgroups r >= R (up to a multiple of 8) mark their slot
// empty.  Workgroup 0 also records the mode in total[1].
__global__ void __launch_bounds__(256) k_tile_order_rx(const uint2* ranges, int R, int T, int n8, uint32_t* order,
                                                       uint32_t* total, int mode)
{
    __shared__ uint32_t cnt[128];
    const int r = blockIdx.x;
    uint32_t* out = order + (size_t)(r % 8) * n8 + (size_t)(r / 8) * T;
    if (r == 0 && threadIdx.x == 0) total[1] = (uint32_t)mode;
    if (r >= R) {
        for (int t = threadIdx.x; t < T; t += 256) out[t] = 0xffffffffu;
        return;
    }
    for (int i = threadIdx.x; i < 128; i += 256) cnt[i] = 0;
    __syncthreads();
    for (int t = threadIdx.x; t < T; t += 256) {
        const uint2 rg = ranges[(size_t)r * T + t];
        atomicAdd(&cnt[length_bucket(rg.y - rg.x)], 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int b = 127; b >= 0; --b) {
            const uint32_t c = cnt[b];
            cnt[b] = run;
            run += c;
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < T; t += 256) {
        const uint2 rg = ranges[(size_t)r * T + t];
        out[atomicAdd(&cnt[length_bucket(rg.y - rg.x)], 1u)] = (uint32_t)(r * T + t);
    }
}

// p[0 .. n) = v: the words up to 16-byte alignment one by one, the bulk as
// 16-byte stores over a grid-stride loop, then the tail.
__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* p, uint32_t v, size_t n)
{
    const size_t head = min(n, (size_t)((16u - ((uintptr_t)p & 15u)) & 15u) / 4);
    const size_t n4 = (n - head) / 4;
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x, nth = (size_t)gridDim.x * blockDim.x;
    if (tid < head) p[tid] = v;
    uint4* q = reinterpret_cast<uint4*>(p + head);
    const uint4 vv = make_uint4(v, v, v, v);
    for (size_t i = tid; i < n4; i += nth) q[i] = vv;
    const size_t t0 = head + 4 * n4;
    if (tid < n - t0) p[t0 + tid] = v;
}

// ---------------------------------------------------------------- launchers
int rs_blocks(int P) { return (P + GS_RS_TILE - 1) / GS_RS_TILE; }

int chunk_len_for(const Dims& d)
{
    // long enough chunks that the count matrix stays within ~R*P entries
    return GS_CHUNK_MIN * max(1, (d.T + GS_CHUNK_MIN - 1) / GS_CHUNK_MIN);
}

int chunk_count_for(const Dims& d) { return max(1, (d.P + chunk_len_for(d) - 1) / chunk_len_for(d)); }

size_t chunk_scatter_lds(const Dims& d) { return GS_SCATTER_WAVES * chunk_scatter_slice(d) * sizeof(uint64_t); }

size_t os_words(const Dims& d)
{
    const size_t nblk = (size_t)rs_blocks(d.P);
    return (size_t)d.R * std::max(nblk * 257, nblk * 256 + GS_OS_GHIST + 4);
}

BinArgs os_args(const Dims& d, const BinArgs& a0)
{
    BinArgs a = a0;
    a.lbstat = a.shist;                                               // [R][nblk][256]
    a.ghist = a.lbstat + (size_t)d.R * rs_blocks(d.P) * 256;          // [R][GS_OS_GHIST]
    a.tick = a.ghist + (size_t)d.R * GS_OS_GHIST;                     // [4][R]
    return a;
}

hipError_t launch_depth_sort(const Dims& d, const BinArgs& a0, hipStream_t s)
{
    if (d.R == 0 || d.P == 0) return hipSuccess;
    const int nblk = rs_blocks(d.P);
    const dim3 grid(nblk, d.R);
    if (nblk > GS_OS_MAX_TILES) {
        // shist as [R][257][nblk] per-tile digit histograms
        for (int pass = 0; pass < 4; ++pass) {
            if (pass == 0) {
                hipLaunchKernelGGL(k_rs_hist<true>, grid, dim3(GS_RS_THREADS), 0, s, d, a0, pass);
                hipLaunchKernelGGL(k_rs_scan, dim3(d.R), dim3(1024), 0, s, d, a0, nblk, 1);
                hipLaunchKernelGGL(k_rs_scatter<true>, grid, dim3(GS_RS_THREADS), 0, s, d, a0, pass);
            } else {
                hipLaunchKernelGGL(k_rs_hist<false>, grid, dim3(GS_RS_THREADS), 0, s, d, a0, pass);
                hipLaunchKernelGGL(k_rs_scan, dim3(d.R), dim3(1024), 0, s, d, a0, nblk, 0);
                hipLaunchKernelGGL(k_rs_scatter<false>, grid, dim3(GS_RS_THREADS), 0, s, d, a0, pass);
            }
        }
        return hipGetLastError();
    }
    const BinArgs a = os_args(d, a0);
    // status words, digit counts and tickets start at zero (a fill kernel:
    // the sort may be captured into a graph)
    hipError_t e = launch_fill_u32(a.shist, 0u, os_words(d), s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_os_hist, grid, dim3(GS_RS_THREADS), 0, s, d, a);
    hipLaunchKernelGGL(k_os_pass<true>, grid, dim3(GS_RS_THREADS), 0, s, d, a, 0, nblk);
    for (int pass = 1; pass < 4; ++pass)
        hipLaunchKernelGGL(k_os_pass<false>, grid, dim3(GS_RS_THREADS), 0, s, d, a, pass, nblk);
    return hipGetLastError();
}

#define GS_CC_CPB 4
hipError_t launch_chunk_count(const Dims& d, const BinArgs& a, hipStream_t s)
{
    if (d.R == 0) return hipSuccess;
    const size_t one = chunk_lds(d) ? (size_t)(d.gx + 1) * d.gy * 4 : 0;
    // four chunks per workgroup where their difference arrays fit in 32 KB
    if (one * GS_CC_CPB <= 32768)
        hipLaunchKernelGGL(k_chunk_count<GS_CC_CPB>, dim3((a.nchunks + GS_CC_CPB - 1) / GS_CC_CPB, d.R),
                           dim3(GS_RS_THREADS), one * GS_CC_CPB, s, d, a);
    else
        hipLaunchKernelGGL(k_chunk_count<1>, dim3(a.nchunks, d.R), dim3(GS_RS_THREADS), one, s, d, a);
    return hipGetLastError();
}

hipError_t launch_chunk_scan(const Dims& d, const BinArgs& a, hipStream_t s)
{
    const int n = d.R * d.T;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_chunk_scan, dim3((n + 3) / 4), dim3(256), 0, s, d, a);
    return hipGetLastError();
}

hipError_t launch_segscan(const Dims& d, const BinArgs& a, hipStream_t s)
{
    const int n = d.R * d.T;
    if (n > 0 && n <= GS_SEGSCAN_LDS)
        hipLaunchKernelGGL(k_segscan_lds, dim3(1), dim3(1024), 0, s, n, a.segcnt, a.segbase, a.ranges, a.total);
    else
        hipLaunchKernelGGL(k_segscan, dim3(1), dim3(1024), 0, s, n, a.segcnt, a.segbase, a.ranges, a.total);
    return hipGetLastError();
}

static_assert((GS_CHUNK_MIN == 256 || GS_CHUNK_MIN == 512) && GS_ROW_PLACE_MAX_T <= 1024,
              "row placement holds 4 to 16 groups of 64");
bool row_place_ok(const Dims& d) { return rect_packed(d) && d.T <= GS_ROW_PLACE_MAX_T; }

template <int NI>
static void launch_row_place_ni(const Dims& d, const BinArgs& a, hipStream_t s)
{
    const dim3 grid((unsigned)((a.nchunks * d.gy + GS_RP_WAVES - 1) / GS_RP_WAVES), (unsigned)d.R);
    const dim3 block(GS_WAVE * GS_RP_WAVES);
    if (a.keys_sorted) hipLaunchKernelGGL((k_row_place<NI, true>), grid, block, 0, s, d, a);
    else hipLaunchKernelGGL((k_row_place<NI, false>), grid, block, 0, s, d, a);
}

hipError_t launch_bin_place(const Dims& d, const BinArgs& a, hipStream_t s)
{
    if (d.R == 0 || d.P == 0) return hipSuccess;
    if (!row_place_ok(d)) return launch_chunk_scatter(d, a, s);
    // chunk_len = GS_CHUNK_MIN * ceil(T / GS_CHUNK_MIN) <= 1024: 4, 8, 12 or 16 groups of 64
    switch (a.chunk_len / GS_WAVE) {
    case 4: launch_row_place_ni<4>(d, a, s); break;
    case 8: launch_row_place_ni<8>(d, a, s); break;
    case 12: launch_row_place_ni<12>(d, a, s); break;
    default: launch_row_place_ni<16>(d, a, s); break;
    }
    return hipGetLastError();
}

hipError_t launch_chunk_scatter(const Dims& d, const BinArgs& a, hipStream_t s)
{
    if (d.R == 0 || d.P == 0) return hipSuccess;
    const unsigned nb = (unsigned)((a.nchunks + GS_SCATTER_WAVES - 1) / GS_SCATTER_WAVES);
    hipLaunchKernelGGL(k_chunk_scatter, dim3(nb, d.R), dim3(GS_WAVE * GS_SCATTER_WAVES), chunk_scatter_lds(d), s,
                       d, a);
    return hipGetLastError();
}

size_t order_length(const Dims& d)
{
    // room for every mode: 8 XCD lists of ceil(R/8) receivers' tiles
    return (size_t)8 * ((d.R + 7) / 8) * d.T;
}

hipError_t launch_segscan_order(const Dims& d, const BinArgs& a, int mode, uint32_t* order, hipStream_t s)
{
    const int n = d.R * d.T;
    if (d.R == 0 || d.T == 0 || n > GS_SEGSCAN_LDS || mode == ORDER_XCD_RECEIVER) {
        const hipError_t e = launch_segscan(d, a, s);
        return e == hipSuccess ? launch_tile_order(d, mode, a.ranges, order, a.total, s) : e;
    }
    const int L = mode == ORDER_XCD_MIXED ? 8 : 1;
    const int stride = mode == ORDER_XCD_MIXED ? (d.R + 7) / 8 * d.T : 0;
    hipLaunchKernelGGL(k_segscan_order, dim3(1 + L), dim3(1024), 0, s, n, a.segcnt, a.segbase, a.ranges, a.total, d.R,
                       d.T, stride, order, mode, (int)order_length(d));
    return hipGetLastError();
}

hipError_t launch_tile_order(const Dims& d, int mode, const uint2* ranges, uint32_t* order, uint32_t* total,
                             hipStream_t s)
{
    if (d.R == 0 || d.T == 0) return hipSuccess;
    // every slot of order_length(d) is written: an entry or the empty marker
    const int n8 = (d.R + 7) / 8 * d.T;
    if (mode == ORDER_XCD_RECEIVER)
        hipLaunchKernelGGL(k_tile_order_rx, dim3((unsigned)((d.R + 7) / 8 * 8)), dim3(256), 0, s, ranges, d.R, d.T, n8,
                           order, total, mode);
    else if (mode == ORDER_XCD_MIXED)
        hipLaunchKernelGGL(k_tile_order, dim3(8), dim3(1024), 0, s, ranges, d.R, d.T, n8, order, total, mode,
                           (int)order_length(d));
    else
        hipLaunchKernelGGL(k_tile_order, dim3(1), dim3(1024), 0, s, ranges, d.R, d.T, 0, order, total, mode,
                           (int)order_length(d));
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((n / 4 + 255) / 256 + 1, 8192);
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

}  // namespace gs
