// Binning for R receivers: per-tile instance counts -> tile ranges, then the
// (tile, depth) ordering of every tile list.  Together these replace
// duplicateWithKeys, the 40-bit radix sort and identifyTileRanges
// (rasterizer_impl.cu:102-171, 270-312) and produce the same ordering:
// within a tile, Gaussians by ascending float depth bits, ties by index.
//
//   k_segscan   exclusive prefix over (receiver, tile) of the per-tile counts
//               (accumulated by the preprocess) -> ranges, K
//   k_scatter   writes (depth bits, index) of every instance into its tile
//               segment: per-block LDS counts, one global atomic per touched
//               tile to reserve space, LDS atomics for the slots
//   k_sort_small one wave per tile segment of up to 2048 instances: LSD radix
//               sort on the depth bits (4-bit digits, ballot ranking) with the
//               segment in registers, no workgroup barriers
//   k_sort_long  one workgroup per longer segment (persistent over the list of
//               long segments): the same sort in LDS when the segment fits,
//               chunk by chunk through global memory otherwise
//   Both then put runs of equal depth in index order (the order the
//   reference's stable sort gives, since its input is in index order).
#include "gs_device.h"
#include "gs_kernels.h"

namespace gs {

// segments up to this length are sorted by one wave (k_sort_small), longer
// ones by a workgroup (k_sort_long)
#define GS_SORT_SMALL 2048

// ---------------------------------------------------------------- scans
// Block-wide exclusive scan of one value per thread (1024 threads max).
template <int NT>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_wave, uint32_t& total)
{
    const int wave = threadIdx.x / GS_WAVE, lane = threadIdx.x % GS_WAVE;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < GS_WAVE; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
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

__global__ void __launch_bounds__(1024) k_segscan(int n, const uint32_t* segcnt, uint32_t* segbase, uint2* ranges,
                                                   uint32_t* total, uint32_t* long_list)
{
    __shared__ uint32_t s_wave[1024 / GS_WAVE];
    const int per = (n + 1023) / 1024;
    const int i0 = min(n, (int)threadIdx.x * per), i1 = min(n, i0 + per);
    uint32_t sum = 0, mx = 0, nlong = 0;
    for (int i = i0; i < i1; ++i) {
        sum += segcnt[i];
        mx = max(mx, segcnt[i]);
        nlong += segcnt[i] > GS_SORT_SMALL ? 1u : 0u;
    }
    uint32_t tot, totlong;
    uint32_t run = block_exclusive_scan<1024>(sum, s_wave, tot);
    uint32_t lpos = block_exclusive_scan<1024>(nlong, s_wave, totlong);
    for (int i = i0; i < i1; ++i) {
        const uint32_t c = segcnt[i];
        segbase[i] = run;
        // empty tiles keep the range (0, 0) as after the reference's memset
        // (rasterizer_impl.cu:303)
        ranges[i] = c ? make_uint2(run, run + c) : make_uint2(0u, 0u);
        if (c > GS_SORT_SMALL) long_list[lpos++] = (uint32_t)i;
        run += c;
    }
    if (threadIdx.x == 0) total[2] = totlong;
    // max segment length
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, off));
    if ((threadIdx.x % GS_WAVE) == 0) s_wave[threadIdx.x / GS_WAVE] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t m = 0;
        for (int w = 0; w < 1024 / GS_WAVE; ++w) m = max(m, s_wave[w]);
        total[0] = tot;
        total[1] = m;
    }
}

// ---------------------------------------------------------------- scatter
__global__ void __launch_bounds__(GS_PRE_THREADS) k_scatter(Dims d, BinArgs a)
{
    __shared__ uint32_t s_cnt[GS_LDS_TILES];
    __shared__ uint32_t s_base[GS_LDS_TILES];
    if ((uint64_t)a.total[0] > a.Kcap) return;    // capacity overflow: reported by the host
    const bool lds_tiles = d.T <= GS_LDS_TILES;
    const int b = blockIdx.x, r = blockIdx.y;
    const int g = b * GS_PRE_THREADS + threadIdx.x;
    const size_t ri = (size_t)r * d.P + g;
    const size_t rT = (size_t)r * d.T;
    int x0 = 0, y0 = 0, x1 = 0, y1 = 0;
    uint32_t dbits = 0;
    if (g < d.P) {
        const int radius = a.radii[ri];
        if (radius > 0) {
            const float4 s0 = a.splat[2 * ri];
            get_rect(s0.x, s0.y, radius, d.gx, d.gy, x0, y0, x1, y1);
            dbits = gs_as_uint(a.depth[ri]);
        }
    }
    if (!lds_tiles) {
        for (int ty = y0; ty < y1; ++ty)
            for (int tx = x0; tx < x1; ++tx) {
                const int t = ty * d.gx + tx;
                const uint32_t pos = a.segbase[rT + t] + atomicAdd(&a.cursor[rT + t], 1u);
                a.keys[pos] = dbits;
                a.vals[pos] = (uint32_t)g;
            }
        return;
    }
    for (int i = threadIdx.x; i < d.T; i += GS_PRE_THREADS) s_cnt[i] = 0;
    __syncthreads();
    for (int ty = y0; ty < y1; ++ty)
        for (int tx = x0; tx < x1; ++tx) atomicAdd(&s_cnt[ty * d.gx + tx], 1u);
    __syncthreads();
    for (int t = threadIdx.x; t < d.T; t += GS_PRE_THREADS) {
        const uint32_t c = s_cnt[t];
        if (c) s_base[t] = a.segbase[rT + t] + atomicAdd(&a.cursor[rT + t], c);
        s_cnt[t] = 0;
    }
    __syncthreads();
    for (int ty = y0; ty < y1; ++ty)
        for (int tx = x0; tx < x1; ++tx) {
            const int t = ty * d.gx + tx;
            const uint32_t pos = s_base[t] + atomicAdd(&s_cnt[t], 1u);
            a.keys[pos] = dbits;
            a.vals[pos] = (uint32_t)g;
        }
}

// ---------------------------------------------------------------- sort
#define GS_SORT_THREADS 512
#define GS_SORT_WAVES (GS_SORT_THREADS / GS_WAVE)
#define GS_SORT_ROWS 16
#define GS_SORT_CHUNK (GS_SORT_THREADS * GS_SORT_ROWS)   // 8192 instances
#define GS_SORT_TABLE (16 * (GS_SORT_ROWS * GS_SORT_WAVES + 1))

struct SortShared {
    uint32_t k[2][GS_SORT_CHUNK];
    uint32_t v[2][GS_SORT_CHUNK];
    uint32_t table[GS_SORT_TABLE];
    uint32_t hist[16];
    uint32_t base[16];
    uint32_t wscan[GS_SORT_WAVES];
    uint32_t red[2][GS_SORT_WAVES];
    uint32_t flag;
};

// Ballot-based stable ranking of 4-bit digits for one row of 512 elements.
__device__ __forceinline__ uint64_t digit_match(bool valid, uint32_t dgt)
{
    uint64_t m = ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 4; ++bit) {
        const bool set = (dgt >> bit) & 1u;
        const uint64_t B = ballot(valid && set);
        m &= set ? B : ~B;
    }
    return m;
}

// One stable LSD pass over n elements (any n), chunk by chunk.
template <typename KP, typename VP, typename DKP, typename DVP>
__device__ void radix_pass(SortShared& sm, int n, KP sk, VP sv, DKP dk, DVP dv, uint32_t kmin, int shift)
{
    const int tid = threadIdx.x, wave = tid / GS_WAVE;
    const uint64_t lt = lanemask_lt();
    const int nchunks = (n + GS_SORT_CHUNK - 1) / GS_SORT_CHUNK;
    if (nchunks > 1) {
        // digit histogram of the whole segment
        if (tid < 16) sm.hist[tid] = 0;
        __syncthreads();
        uint32_t pk[4] = {0, 0, 0, 0};   // 16 packed 8-bit counters... flushed every 255
        int cnt = 0;
        for (int e = tid; e < n; e += GS_SORT_THREADS) {
            const uint32_t dg = ((sk[e] - kmin) >> shift) & 15u;
            pk[dg >> 2] += 1u << (8 * (dg & 3));
            if (++cnt == 255) {
                for (int q = 0; q < 16; ++q) {
                    const uint32_t c = (pk[q >> 2] >> (8 * (q & 3))) & 255u;
                    if (c) atomicAdd(&sm.hist[q], c);
                }
                pk[0] = pk[1] = pk[2] = pk[3] = 0;
                cnt = 0;
            }
        }
        for (int q = 0; q < 16; ++q) {
            const uint32_t c = (pk[q >> 2] >> (8 * (q & 3))) & 255u;
            if (c) atomicAdd(&sm.hist[q], c);
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t run = 0;
            for (int q = 0; q < 16; ++q) { sm.base[q] = run; run += sm.hist[q]; }
        }
        __syncthreads();
    }
    for (int ch = 0; ch < nchunks; ++ch) {
        const int e0 = ch * GS_SORT_CHUNK;
        const int cn = min(GS_SORT_CHUNK, n - e0);
        const int rows = (cn + GS_SORT_THREADS - 1) / GS_SORT_THREADS;
        const int DS = rows * GS_SORT_WAVES + 1;   // digit stride, padded against bank conflicts
        const int S = 16 * DS;
        for (int i = tid; i < S; i += GS_SORT_THREADS) sm.table[i] = 0;
        __syncthreads();
        for (int i = 0; i < rows; ++i) {
            const int e = i * GS_SORT_THREADS + tid;
            const bool valid = e < cn;
            const uint32_t dg = valid ? ((sk[e0 + e] - kmin) >> shift) & 15u : 0u;
            const uint64_t m = digit_match(valid, dg);
            if (valid && (m & lt) == 0)
                sm.table[dg * DS + i * GS_SORT_WAVES + wave] = (uint32_t)__popcll(m);
        }
        __syncthreads();
        // exclusive scan of the table in (digit, row, wave) order
        {
            const int per = (S + GS_SORT_THREADS - 1) / GS_SORT_THREADS;
            const int i0 = min(S, tid * per), i1 = min(S, i0 + per);
            uint32_t s = 0;
            for (int i = i0; i < i1; ++i) s += sm.table[i];
            uint32_t tot;
            uint32_t run = block_exclusive_scan<GS_SORT_THREADS>(s, sm.wscan, tot);
            for (int i = i0; i < i1; ++i) {
                const uint32_t c = sm.table[i];
                sm.table[i] = run;
                run += c;
            }
        }
        __syncthreads();
        // per-digit start inside this chunk: P(d,0,0)
        for (int i = 0; i < rows; ++i) {
            const int e = i * GS_SORT_THREADS + tid;
            const bool valid = e < cn;
            uint32_t key = 0;
            uint32_t dg = 0;
            if (valid) {
                key = sk[e0 + e];
                dg = ((key - kmin) >> shift) & 15u;
            }
            const uint64_t m = digit_match(valid, dg);
            if (valid) {
                uint32_t pos = sm.table[dg * DS + i * GS_SORT_WAVES + wave] + (uint32_t)__popcll(m & lt);
                if (nchunks > 1) pos = pos - sm.table[dg * DS] + sm.base[dg];
                dk[pos] = key;
                dv[pos] = sv[e0 + e];
            }
        }
        __syncthreads();
        if (nchunks > 1) {
            if (tid < 16) {
                const uint32_t start = sm.table[tid * DS];
                const uint32_t end = (tid < 15) ? sm.table[(tid + 1) * DS] : (uint32_t)cn;
                sm.base[tid] += end - start;
            }
            __syncthreads();
        }
    }
}

// Runs of equal depth bits come out of the radix passes in scatter order;
// put each run in ascending Gaussian index, the order the reference's stable
// sort produces from its index-ordered input.  Runs are rare and short.
template <typename KP, typename VP>
__device__ void order_ties(int n, KP k, VP v, uint32_t* s_flag)
{
    const int tid = threadIdx.x;
    if (tid == 0) *s_flag = 0;
    __syncthreads();
    bool any = false;
    for (int e = tid + 1; e < n; e += GS_SORT_THREADS) any |= (k[e] == k[e - 1]);
    if (any) *s_flag = 1;
    __syncthreads();
    if (*s_flag) {
        for (int e = tid; e < n; e += GS_SORT_THREADS) {
            const bool starts = (e == 0 || k[e - 1] != k[e]) && e + 1 < n && k[e + 1] == k[e];
            if (!starts) continue;
            int end = e + 1;
            while (end < n && k[end] == k[e]) ++end;
            for (int i = e + 1; i < end; ++i) {        // insertion sort by index
                const uint32_t x = v[i];
                int j = i - 1;
                while (j >= e && v[j] > x) { v[j + 1] = v[j]; --j; }
                v[j + 1] = x;
            }
        }
    }
    __syncthreads();
}

// One workgroup per long segment (> GS_SORT_SMALL instances).  Persistent:
// the grid walks the compacted list of long segments written by k_segscan.
__global__ void __launch_bounds__(GS_SORT_THREADS) k_sort_long(BinArgs a)
{
    extern __shared__ unsigned char s_raw[];
    SortShared& sm = *reinterpret_cast<SortShared*>(s_raw);
    if ((uint64_t)a.total[0] > a.Kcap) return;
    const int tid = threadIdx.x;
    const uint32_t nlong = a.total[2];
    for (uint32_t li = blockIdx.x; li < nlong; li += gridDim.x) {
        const uint32_t seg = a.long_list[li];
        const int n = (int)a.ranges[seg].y - (int)a.ranges[seg].x;
        const uint32_t base = a.ranges[seg].x;
        uint32_t* gk = a.keys + base;
        uint32_t* gv = a.vals + base;
        // key range of the segment
        uint32_t kmin = 0xffffffffu, kmax = 0;
        for (int e = tid; e < n; e += GS_SORT_THREADS) {
            const uint32_t k = gk[e];
            kmin = min(kmin, k);
            kmax = max(kmax, k);
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            kmin = min(kmin, (uint32_t)__shfl_xor(kmin, off));
            kmax = max(kmax, (uint32_t)__shfl_xor(kmax, off));
        }
        if ((tid % GS_WAVE) == 0) { sm.red[0][tid / GS_WAVE] = kmin; sm.red[1][tid / GS_WAVE] = kmax; }
        __syncthreads();
        kmin = sm.red[0][0];
        kmax = sm.red[1][0];
        for (int w = 1; w < GS_SORT_WAVES; ++w) { kmin = min(kmin, sm.red[0][w]); kmax = max(kmax, sm.red[1][w]); }
        const uint32_t span = kmax - kmin;
        const int nbits = span ? 32 - __clz((int)span) : 0;
        const int passes = (nbits + 3) / 4;
        __syncthreads();

        if (n <= GS_SORT_CHUNK) {
            for (int e = tid; e < n; e += GS_SORT_THREADS) { sm.k[0][e] = gk[e]; sm.v[0][e] = gv[e]; }
            __syncthreads();
            int cur = 0;
            for (int p = 0; p < passes; ++p) {
                radix_pass(sm, n, sm.k[cur], sm.v[cur], sm.k[cur ^ 1], sm.v[cur ^ 1], kmin, 4 * p);
                cur ^= 1;
            }
            order_ties(n, sm.k[cur], sm.v[cur], &sm.flag);
            for (int e = tid; e < n; e += GS_SORT_THREADS) {
                a.plist[base + e] = sm.v[cur][e];
                if (a.keys_sorted) a.keys_sorted[base + e] = sm.k[cur][e];
            }
        } else {
            uint32_t* tk = a.keys_tmp + base;
            uint32_t* tv = a.vals_tmp + base;
            int cur = 0;
            for (int p = 0; p < passes; ++p) {
                if (cur == 0) radix_pass(sm, n, gk, gv, tk, tv, kmin, 4 * p);
                else radix_pass(sm, n, tk, tv, gk, gv, kmin, 4 * p);
                cur ^= 1;
            }
            uint32_t* fk = cur ? tk : gk;
            uint32_t* fv = cur ? tv : gv;
            order_ties(n, fk, fv, &sm.flag);
            for (int e = tid; e < n; e += GS_SORT_THREADS) {
                a.plist[base + e] = fv[e];
                if (a.keys_sorted) a.keys_sorted[base + e] = fk[e];
            }
        }
        __syncthreads();
    }
}

// One wave per short segment (<= GS_SORT_SMALL instances), no workgroup
// barriers.  The segment is double-buffered in LDS (2 x 16 KiB).  Each 4-bit
// LSD pass takes four ballots per row of 64: lane d < 16 derives from them how
// many elements of the row carry digit d (so the digit histogram and the
// running digit offsets stay in lanes 0..15, with no LDS read-modify-write
// chain) and every element derives its stable rank inside the row.  Runs of
// equal depth are finally put in index order.
__device__ __forceinline__ uint64_t digit_mask_of(uint32_t dg, uint64_t V, const uint64_t* B)
{
    uint64_t m = V;
#pragma unroll
    for (int bit = 0; bit < 4; ++bit) m &= ((dg >> bit) & 1u) ? B[bit] : ~B[bit];
    return m;
}

__global__ void __launch_bounds__(GS_WAVE) k_sort_small(BinArgs a, int nseg)
{
    __shared__ uint2 buf[2][GS_SORT_SMALL];
    if ((uint64_t)a.total[0] > a.Kcap) return;
    const int seg = blockIdx.x;
    if (seg >= nseg) return;
    const uint2 rg = a.ranges[seg];
    const int n = (int)(rg.y - rg.x);
    if (n <= 0 || n > GS_SORT_SMALL) return;
    const int lane = threadIdx.x;
    const uint64_t lt = lanemask_lt();
    const int rows = (n + GS_WAVE - 1) / GS_WAVE;
    uint32_t kmin = 0xffffffffu, kmax = 0;
    for (int e = lane; e < n; e += GS_WAVE) {
        const uint2 t = make_uint2(a.keys[rg.x + e], a.vals[rg.x + e]);
        buf[0][e] = t;
        kmin = min(kmin, t.x);
        kmax = max(kmax, t.x);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        kmin = min(kmin, (uint32_t)__shfl_xor(kmin, off));
        kmax = max(kmax, (uint32_t)__shfl_xor(kmax, off));
    }
    const uint32_t span = kmax - kmin;
    const int nbits = span ? 32 - __clz((int)span) : 0;
    const int passes = (nbits + 3) / 4;
    const uint32_t my_digit = (uint32_t)(lane & 15);
    __syncthreads();
    int cur = 0;
    for (int p = 0; p < passes; ++p) {
        const int shift = 4 * p;
        const uint2* src = buf[cur];
        uint2* dst = buf[cur ^ 1];
        // digit histogram: lane d < 16 counts digit d
        uint32_t cnt = 0;
        for (int i = 0; i < rows; ++i) {
            const int e = i * GS_WAVE + lane;
            const bool valid = e < n;
            const uint32_t dg = valid ? ((src[e].x - kmin) >> shift) & 15u : 0u;
            uint64_t B[4];
#pragma unroll
            for (int bit = 0; bit < 4; ++bit) B[bit] = ballot(valid && ((dg >> bit) & 1u));
            cnt += (uint32_t)__popcll(digit_mask_of(my_digit, ballot(valid), B));
        }
        // exclusive scan over the 16 digits
        uint32_t x = lane < 16 ? cnt : 0u;
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        uint32_t base = x - (lane < 16 ? cnt : 0u);
        // stable scatter
        for (int i = 0; i < rows; ++i) {
            const int e = i * GS_WAVE + lane;
            const bool valid = e < n;
            const uint2 t = valid ? src[e] : make_uint2(0u, 0u);
            const uint32_t dg = ((t.x - kmin) >> shift) & 15u;
            uint64_t B[4];
#pragma unroll
            for (int bit = 0; bit < 4; ++bit) B[bit] = ballot(valid && ((dg >> bit) & 1u));
            const uint64_t V = ballot(valid);
            const uint32_t rank = (uint32_t)__popcll(digit_mask_of(dg, V, B) & lt);
            const uint32_t b = (uint32_t)__shfl((int)base, (int)dg);
            if (valid) dst[b + rank] = t;
            base += (uint32_t)__popcll(digit_mask_of(my_digit, V, B));
        }
        __syncthreads();
        cur ^= 1;
    }
    // runs of equal depth bits -> index order
    uint2* fin = buf[cur];
    for (int e = lane; e < n; e += GS_WAVE) {
        const bool starts = (e == 0 || fin[e - 1].x != fin[e].x) && e + 1 < n && fin[e + 1].x == fin[e].x;
        if (!starts) continue;
        int end = e + 1;
        while (end < n && fin[end].x == fin[e].x) ++end;
        for (int i = e + 1; i < end; ++i) {
            const uint2 t = fin[i];
            int j = i - 1;
            while (j >= e && fin[j].y > t.y) { fin[j + 1] = fin[j]; --j; }
            fin[j + 1] = t;
        }
    }
    __syncthreads();
    for (int e = lane; e < n; e += GS_WAVE) {
        const uint2 t = fin[e];
        a.plist[rg.x + e] = t.y;
        if (a.keys_sorted) a.keys_sorted[rg.x + e] = t.x;
    }
}

hipError_t launch_segscan(const Dims& d, const BinArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(k_segscan, dim3(1), dim3(1024), 0, s, d.R * d.T, a.segcnt, a.segbase, a.ranges, a.total,
                       a.long_list);
    return hipGetLastError();
}

hipError_t launch_scatter(const Dims& d, const BinArgs& a, hipStream_t s)
{
    if (d.NB > 0 && d.R > 0)
        hipLaunchKernelGGL(k_scatter, dim3(d.NB, d.R), dim3(GS_PRE_THREADS), 0, s, d, a);
    return hipGetLastError();
}

hipError_t launch_sort_small(const Dims& d, const BinArgs& a, hipStream_t s)
{
    const int nseg = d.R * d.T;
    if (nseg == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sort_small, dim3(nseg), dim3(GS_WAVE), 0, s, a, nseg);
    return hipGetLastError();
}

hipError_t launch_sort_long(const Dims& d, const BinArgs& a, hipStream_t s)
{
    const int nseg = d.R * d.T;
    if (nseg == 0) return hipSuccess;
    // stages up to GS_SORT_CHUNK instances in LDS (about 136 KiB of the
    // 160 KiB a gfx950 workgroup can hold); one block per CU
    static hipError_t attr = hipFuncSetAttribute((const void*)k_sort_long, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)sizeof(SortShared));
    if (attr != hipSuccess) return attr;
    hipLaunchKernelGGL(k_sort_long, dim3(std::min(nseg, 256)), dim3(GS_SORT_THREADS), sizeof(SortShared), s, a);
    return hipGetLastError();
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

}  // namespace gs
