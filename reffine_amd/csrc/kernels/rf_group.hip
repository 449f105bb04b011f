// Ahead-of-time kernels of the hash-join + grouped-aggregation path: key
// range, the rank map, a stable LSD radix sort with payload columns,
// adjacent-duplicate detection, segment bounds of a sorted group column, the
// partitioning of joined rows by group rank, and the union merge's tile
// splits.
//
// The deterministic mode orders joined records by group rank with a stable
// sort (ties keep ascending probe-row order) and folds each group
// sequentially: the fold order of the reference's nested loop (an outer Op
// over the build table whose output reduces the probe rows with a matching
// key, src/pass/loopgen.cpp:206-269). The default mode partitions the rows by
// rank range and folds each partition with LDS atomics (generated
// rf_gj_agg). Neither uses random global atomics, which saturate near
// 23 G updates/s on gfx950.
#include <hip/hip_runtime.h>

#include "rf_device.h"

#define RF_SORT_BLOCK 256
#define RF_SORT_ITEMS 16
#define RF_SORT_TILE (RF_SORT_BLOCK * RF_SORT_ITEMS)
#define RF_SORT_MAXPAY 8

struct RfSortPass {
    const void* kin;
    void* kout;
    const void* pin[RF_SORT_MAXPAY];
    void* pout[RF_SORT_MAXPAY];
    int32_t pw[RF_SORT_MAXPAY];  // payload widths in bytes (1, 2, 4, 8)
    int32_t np;
    int32_t shift;               // digit = ((key - bias) >> shift) & 255
    uint64_t bias;
    int64_t n;
    int64_t nblocks;
    uint32_t* hist;              // [256][nblocks], digit-major; scanned in place
};

template <typename K>
__device__ __forceinline__ uint32_t rf_digit(const RfSortPass& P, int64_t e) {
    const uint64_t k = (uint64_t)((const K*)P.kin)[e];
    return (uint32_t)((k - P.bias) >> P.shift) & 255u;
}

template <typename K>
__device__ __forceinline__ void rf_sort_hist(const RfSortPass& P) {
    __shared__ uint32_t h[256];
    const int t = threadIdx.x;
    h[t] = 0;
    __syncthreads();
    const int64_t e0 = (int64_t)blockIdx.x * RF_SORT_TILE;
    const int64_t e1 = e0 + RF_SORT_TILE < P.n ? e0 + RF_SORT_TILE : P.n;
    for (int64_t e = e0 + t; e < e1; e += RF_SORT_BLOCK) atomicAdd(&h[rf_digit<K>(P, e)], 1u);
    __syncthreads();
    P.hist[(int64_t)t * P.nblocks + blockIdx.x] = h[t];
}

// The same histogram with 2^RB digits (RB = 10: two passes where 8-bit
// digits need three).
template <typename K, int RB>
__device__ __forceinline__ void rf_sort_hist_w(const RfSortPass& P) {
    constexpr int R = 1 << RB;
    __shared__ uint32_t h[R];
    const int t = threadIdx.x;
    for (int d = t; d < R; d += RF_SORT_BLOCK) h[d] = 0;
    __syncthreads();
    const int64_t e0 = (int64_t)blockIdx.x * RF_SORT_TILE;
    const int64_t e1 = e0 + RF_SORT_TILE < P.n ? e0 + RF_SORT_TILE : P.n;
    for (int64_t e = e0 + t; e < e1; e += RF_SORT_BLOCK) {
        const uint64_t k = (uint64_t)((const K*)P.kin)[e];
        atomicAdd(&h[(uint32_t)((k - P.bias) >> P.shift) & (uint32_t)(R - 1)], 1u);
    }
    __syncthreads();
    for (int d = t; d < R; d += RF_SORT_BLOCK) P.hist[(int64_t)d * P.nblocks + blockIdx.x] = h[d];
}

// Stable scatter of one tile. Element e = tile0 + round*256 + thread; its
// destination is the digit's running offset for this block, plus the
// elements of the same digit in lower waves of the round, plus those in
// lower lanes of its wave (found with eight ballots: the lanes whose digit
// matches bit by bit).
template <typename K>
__device__ __forceinline__ void rf_sort_scatter(const RfSortPass& P) {
    __shared__ uint32_t goff[256];
    __shared__ uint32_t cnt[RF_SORT_BLOCK / 64][256];
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    goff[t] = P.hist[(int64_t)t * P.nblocks + blockIdx.x];
#pragma unroll
    for (int v = 0; v < RF_SORT_BLOCK / 64; v++) cnt[v][t] = 0;
    __syncthreads();
    const int64_t e0 = (int64_t)blockIdx.x * RF_SORT_TILE;
    const int64_t e1 = e0 + RF_SORT_TILE < P.n ? e0 + RF_SORT_TILE : P.n;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int64_t base = e0; base < e1; base += RF_SORT_BLOCK) {
        const int64_t e = base + t;
        const bool ok = e < e1;
        const uint32_t d = ok ? rf_digit<K>(P, e) : 0u;
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool on = (d >> b) & 1u;
            const uint64_t m = __ballot(on);
            peers &= on ? m : ~m;
        }
        const uint32_t below = (uint32_t)__popcll(peers & lt);
        if (ok && below == 0) cnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (ok) {
            uint32_t pos = goff[d] + below;
            for (int v = 0; v < w; v++) pos += cnt[v][d];
            ((K*)P.kout)[pos] = ((const K*)P.kin)[e];
            for (int j = 0; j < P.np; j++) {
                switch (P.pw[j]) {
                    case 1: ((uint8_t*)P.pout[j])[pos] = ((const uint8_t*)P.pin[j])[e]; break;
                    case 2: ((uint16_t*)P.pout[j])[pos] = ((const uint16_t*)P.pin[j])[e]; break;
                    case 4: ((uint32_t*)P.pout[j])[pos] = ((const uint32_t*)P.pin[j])[e]; break;
                    default: ((uint64_t*)P.pout[j])[pos] = ((const uint64_t*)P.pin[j])[e]; break;
                }
            }
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int v = 0; v < RF_SORT_BLOCK / 64; v++) {
            add += cnt[v][t];
            cnt[v][t] = 0;
        }
        goff[t] += add;
        __syncthreads();
    }
}

template <typename T>
__device__ __forceinline__ void rf_copy_elem(void* dst, int64_t di, const void* src, int64_t si) {
    ((T*)dst)[di] = ((const T*)src)[si];
}

__device__ __forceinline__ void rf_copy_w(int w, void* dst, int64_t di, const void* src, int64_t si) {
    switch (w) {
        case 1: rf_copy_elem<uint8_t>(dst, di, src, si); break;
        case 2: rf_copy_elem<uint16_t>(dst, di, src, si); break;
        case 4: rf_copy_elem<uint32_t>(dst, di, src, si); break;
        default: rf_copy_elem<uint64_t>(dst, di, src, si); break;
    }
}

// One payload column of a wave's elements into its LDS plane at their local
// positions (all of this lane's loads in flight together).
template <typename T>
__device__ __forceinline__ void rf_place_plane(void* plane, const uint32_t (&lpos)[RF_SORT_ITEMS],
                                               const void* src, int64_t wbase, int64_t e1) {
    const int lane = threadIdx.x & 63;
    T v[RF_SORT_ITEMS];
#pragma unroll
    for (int r = 0; r < RF_SORT_ITEMS; r++) {
        const int64_t e = wbase + r * 64 + lane;
        v[r] = e < e1 ? ((const T*)src)[e] : (T)0;
    }
#pragma unroll
    for (int r = 0; r < RF_SORT_ITEMS; r++)
        if (wbase + r * 64 + lane < e1) ((T*)plane)[lpos[r]] = v[r];
}

// The same stable scatter, staged through LDS. Each wave ranks its own
// contiguous quarter of the tile against per-wave digit counters (ballots
// find the lanes with the same digit; the wave's lowest such lane advances
// the counter, read by all of them first -- a wave's LDS operations are in
// order, so no barrier is needed inside the loop); a prefix over the waves
// per digit and the tile's digit starts then give every element its place
// in the tile in digit order. The tile is placed in LDS and written out in
// LDS order, so consecutive threads write consecutive positions of a
// digit's output range. The tile's count of digit d is the difference of
// consecutive entries of the scanned (digit-major) histogram. Dynamic LDS:
// TILE keys, then one plane of TILE values per payload.
template <typename K, int RB>
__device__ __forceinline__ void rf_sort_scatter_lds(const RfSortPass& P) {
    constexpr int R = 1 << RB;          // digits
    constexpr int DT = R / RF_SORT_BLOCK;  // digits per thread
    constexpr int WAVES = RF_SORT_BLOCK / 64;
    constexpr int64_t WTILE = RF_SORT_TILE / WAVES;  // elements per wave
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t goff[R];
    __shared__ uint16_t lstart[R];      // <= RF_SORT_TILE
    __shared__ uint16_t cnt[WAVES][R];  // <= WTILE
    __shared__ uint32_t s_scan[RF_SORT_BLOCK];
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    const int64_t e0 = (int64_t)blockIdx.x * RF_SORT_TILE;
    const int64_t e1 = e0 + RF_SORT_TILE < P.n ? e0 + RF_SORT_TILE : P.n;
    const int64_t m = (int64_t)R * P.nblocks;
    // Thread t owns digits [t*DT, (t+1)*DT): their global offsets, their
    // counts in this tile (differences of the scanned histogram) and their
    // local starts (an exclusive scan over the digits).
    uint32_t c[DT];
    uint32_t csum = 0;
#pragma unroll
    for (int j = 0; j < DT; j++) {
        const int d = t * DT + j;
        const int64_t idx = (int64_t)d * P.nblocks + blockIdx.x;
        const uint32_t g = P.hist[idx];
        c[j] = (idx + 1 < m ? P.hist[idx + 1] : (uint32_t)P.n) - g;
        goff[d] = g;
        csum += c[j];
#pragma unroll
        for (int v = 0; v < WAVES; v++) cnt[v][d] = 0;
    }
    s_scan[t] = csum;
    __syncthreads();
    for (int off = 1; off < RF_SORT_BLOCK; off <<= 1) {
        const uint32_t x = t >= off ? s_scan[t - off] : 0u;
        __syncthreads();
        s_scan[t] += x;
        __syncthreads();
    }
    {
        uint32_t run = s_scan[t] - csum;
#pragma unroll
        for (int j = 0; j < DT; j++) {
            lstart[t * DT + j] = (uint16_t)run;
            run += c[j];
        }
    }
    // Ranking: each wave over its contiguous quarter of the tile (see above).
    const int64_t wbase = e0 + (int64_t)w * WTILE;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    K kr[RF_SORT_ITEMS];
    uint32_t lpos[RF_SORT_ITEMS];
    uint32_t dig[RF_SORT_ITEMS];
#pragma unroll
    for (int r = 0; r < RF_SORT_ITEMS; r++) {
        const int64_t e = wbase + r * 64 + lane;
        kr[r] = e < e1 ? ((const K*)P.kin)[e] : (K)0;
    }
#pragma unroll
    for (int r = 0; r < RF_SORT_ITEMS; r++) {
        const bool ok = wbase + r * 64 + lane < e1;
        const uint32_t d =
            ok ? (uint32_t)(((uint64_t)kr[r] - P.bias) >> P.shift) & (uint32_t)(R - 1) : 0u;
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < RB; b++) {
            const bool on = (d >> b) & 1u;
            const uint64_t mk = __ballot(on);
            peers &= on ? mk : ~mk;
        }
        const uint32_t below = (uint32_t)__popcll(peers & lt);
        const uint32_t before = cnt[w][d];
        if (ok && below == 0) cnt[w][d] = (uint16_t)(before + (uint32_t)__popcll(peers));
        lpos[r] = before + below;  // rank among the wave's elements of digit d
        dig[r] = d;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < DT; j++) {
        // Per digit: exclusive prefix over the waves, from the tile start.
        const int d = t * DT + j;
        uint32_t run = lstart[d];
#pragma unroll
        for (int v = 0; v < WAVES; v++) {
            const uint32_t x = cnt[v][d];
            cnt[v][d] = (uint16_t)run;
            run += x;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RF_SORT_ITEMS; r++) lpos[r] += cnt[w][dig[r]];
    K* sk = (K*)lds;
    unsigned char* planes = lds + (size_t)RF_SORT_TILE * sizeof(K);
#pragma unroll
    for (int r = 0; r < RF_SORT_ITEMS; r++)
        if (wbase + r * 64 + lane < e1) sk[lpos[r]] = kr[r];
    size_t poff = 0;
    for (int j = 0; j < P.np; j++) {
        void* plane = planes + poff;
        switch (P.pw[j]) {
            case 1: rf_place_plane<uint8_t>(plane, lpos, P.pin[j], wbase, e1); break;
            case 2: rf_place_plane<uint16_t>(plane, lpos, P.pin[j], wbase, e1); break;
            case 4: rf_place_plane<uint32_t>(plane, lpos, P.pin[j], wbase, e1); break;
            default: rf_place_plane<uint64_t>(plane, lpos, P.pin[j], wbase, e1); break;
        }
        poff += (size_t)RF_SORT_TILE * P.pw[j];
    }
    __syncthreads();
    const int tn = (int)(e1 - e0);
    for (int i = t; i < tn; i += RF_SORT_BLOCK) {
        const K key = sk[i];
        const uint32_t d = (uint32_t)(((uint64_t)key - P.bias) >> P.shift) & (uint32_t)(R - 1);
        const int64_t pos = (int64_t)goff[d] + (i - (int64_t)lstart[d]);
        ((K*)P.kout)[pos] = key;
        size_t off = 0;
        for (int j = 0; j < P.np; j++) {
            rf_copy_w(P.pw[j], P.pout[j], pos, planes + off, i);
            off += (size_t)RF_SORT_TILE * P.pw[j];
        }
    }
}
extern "C" __global__ __launch_bounds__(RF_SORT_BLOCK) void rf_sort_hist_u32(RfSortPass P) {
    rf_sort_hist<uint32_t>(P);
}
extern "C" __global__ __launch_bounds__(RF_SORT_BLOCK) void rf_sort_hist_u64(RfSortPass P) {
    rf_sort_hist<uint64_t>(P);
}
extern "C" __global__ __launch_bounds__(RF_SORT_BLOCK) void rf_sort_scatter_u32(RfSortPass P) {
    rf_sort_scatter<uint32_t>(P);
}
extern "C" __global__ __launch_bounds__(RF_SORT_BLOCK) void rf_sort_scatter_u64(RfSortPass P) {
    rf_sort_scatter<uint64_t>(P);
}
extern "C" __global__ __launch_bounds__(RF_SORT_BLOCK) void rf_sort_scatter_lds_u32(RfSortPass P) {
    rf_sort_scatter_lds<uint32_t, 8>(P);
}
extern "C" __global__ __launch_bounds__(RF_SORT_BLOCK) void rf_sort_scatter_lds_u64(RfSortPass P) {
    rf_sort_scatter_lds<uint64_t, 8>(P);
}
extern "C" __global__ __launch_bounds__(RF_SORT_BLOCK) void rf_sort_hist10_u32(RfSortPass P) {
    rf_sort_hist_w<uint32_t, 10>(P);
}
extern "C" __global__ __launch_bounds__(RF_SORT_BLOCK) void rf_sort_hist10_u64(RfSortPass P) {
    rf_sort_hist_w<uint64_t, 10>(P);
}
extern "C" __global__ __launch_bounds__(RF_SORT_BLOCK) void rf_sort_scatter10_u32(RfSortPass P) {
    rf_sort_scatter_lds<uint32_t, 10>(P);
}
extern "C" __global__ __launch_bounds__(RF_SORT_BLOCK) void rf_sort_scatter10_u64(RfSortPass P) {
    rf_sort_scatter_lds<uint64_t, 10>(P);
}

// Exclusive scan of m histogram entries in three launches: per-chunk sums,
// a one-block scan of the chunk sums, per-chunk scans with the offset.
#define RF_SCAN_CHUNK 4096

__device__ __forceinline__ uint32_t rf_block_sum_u32(uint32_t v, uint32_t* s) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += s[w];
    return t;
}

extern "C" __global__ __launch_bounds__(256) void rf_scan_chunks(const uint32_t* __restrict__ h,
                                                                 int64_t m, uint32_t* part) {
    __shared__ uint32_t s[4];
    const int64_t a = (int64_t)blockIdx.x * RF_SCAN_CHUNK;
    const int64_t b = a + RF_SCAN_CHUNK < m ? a + RF_SCAN_CHUNK : m;
    uint32_t v = 0;
    for (int64_t i = a + threadIdx.x; i < b; i += 256) v += h[i];
    v = rf_block_sum_u32(v, s);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

extern "C" __global__ __launch_bounds__(256) void rf_scan_apply(uint32_t* h, int64_t m,
                                                                const uint32_t* __restrict__ part) {
    __shared__ uint32_t s[256];
    const int t = threadIdx.x;
    const int64_t a = (int64_t)blockIdx.x * RF_SCAN_CHUNK + (int64_t)t * 16;
    uint32_t v[16];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        v[k] = (a + k < m) ? h[a + k] : 0u;
        sum += v[k];
    }
    s[t] = sum;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t x = t >= off ? s[t - off] : 0u;
        __syncthreads();
        s[t] += x;
        __syncthreads();
    }
    uint32_t run = part[blockIdx.x] + s[t] - sum;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        if (a + k < m) h[a + k] = run;
        run += v[k];
    }
}

// Exclusive scan in place by one block of 1024 threads, each owning a
// contiguous range (the chunk sums of the scan above).
extern "C" __global__ __launch_bounds__(1024) void rf_sort_scan(uint32_t* hist, int64_t m) {
    __shared__ uint32_t s[1024];
    const int t = threadIdx.x;
    const int64_t per = (m + 1023) / 1024;
    const int64_t a = (int64_t)t * per;
    const int64_t b = a + per < m ? a + per : m;
    uint32_t sum = 0;
    for (int64_t i = a; i < b; i++) sum += hist[i];
    s[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint32_t v = t >= off ? s[t - off] : 0u;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    uint32_t run = s[t] - sum;
    for (int64_t i = a; i < b; i++) {
        const uint32_t v = hist[i];
        hist[i] = run;
        run += v;
    }
}

// Minimum and maximum of n int64 keys (n read from device memory), as
// sign-flipped unsigned values: out[0] = min ^ 2^63, out[1] = max ^ 2^63.
// out must hold {~0, 0} before the launch. One atomic pair per block (a
// launch of at most 256 blocks): atomics on one address serialise.
extern "C" __global__ __launch_bounds__(256) void rf_minmax_i64(const int64_t* __restrict__ keys,
                                                                const int64_t* __restrict__ n_dev,
                                                                unsigned long long* out) {
    __shared__ unsigned long long smn[4], smx[4];
    const int64_t n = *n_dev;
    unsigned long long mn = ~0ull, mx = 0ull;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long u = (unsigned long long)keys[i] ^ 0x8000000000000000ull;
        mn = u < mn ? u : mn;
        mx = u > mx ? u : mx;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long a = __shfl_xor(mn, o), b = __shfl_xor(mx, o);
        mn = a < mn ? a : mn;
        mx = b > mx ? b : mx;
    }
    if ((threadIdx.x & 63) == 0) {
        smn[threadIdx.x >> 6] = mn;
        smx[threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; w++) {
            mn = smn[w] < mn ? smn[w] : mn;
            mx = smx[w] > mx ? smx[w] : mx;
        }
        atomicMin(&out[0], mn);
        atomicMax(&out[1], mx);
    }
}

// Bitmap filter of a key set: bit (hash >> 32) & fmask of every key (the
// slot index uses the low bits of the same hash).
extern "C" __global__ void rf_filter_build(const int64_t* __restrict__ keys, int64_t n,
                                           uint32_t* filter, uint64_t fmask) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t fb = (rf_hash64((uint64_t)keys[i]) >> 32) & fmask;
        atomicOr(&filter[fb >> 5], 1u << (fb & 31));
    }
}

// Sets *flag when two adjacent sorted keys are equal.
extern "C" __global__ void rf_adjacent_dups(const int64_t* __restrict__ keys, int64_t n,
                                            unsigned int* flag) {
    for (int64_t i = 1 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        if (keys[i] == keys[i - 1]) atomicOr(flag, 1u);
}

// Segment bounds of a sorted rank column: rows [lo[r], hi[r]) hold rank r.
// lo and hi must be zero for ranks without rows.
extern "C" __global__ void rf_seg_bounds(const uint32_t* __restrict__ rank, int64_t n,
                                         int64_t* lo, int64_t* hi) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t r = rank[i];
        if (i == 0 || rank[i - 1] != r) lo[r] = i;
        if (i == n - 1 || rank[i + 1] != r) hi[r] = i + 1;
    }
}

// Rank map of sorted distinct keys: mark bit (key - base); then the popcount
// of each 64-bit word (scanned afterwards into the words' first ranks).
// Lanes of a wave whose keys fall in the same 64-bit word combine their bits
// first (sorted keys put a wave's 64 keys in one or two words), so one lane
// per word issues the atomic. The loop steps whole waves: every lane takes
// part in the shuffles.
// With n_dev set, the keys are the first min(*n_dev, n) (a count an earlier
// kernel wrote). A key outside the map's `words` words marks nothing (it
// cannot happen when the range came from the keys; the caller's count of
// marked keys then falls short of the keys' and reports it).
extern "C" __global__ void rf_rank_bits(const int64_t* __restrict__ keys, int64_t n,
                                        const int64_t* __restrict__ n_dev, int64_t base,
                                        unsigned long long* bits, int64_t words) {
    if (n_dev && *n_dev < n) n = *n_dev;
    const int lane = threadIdx.x & 63;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); i0 < n;
         i0 += stride) {
        const int64_t i = i0 + lane;
        const uint64_t off0 = i < n ? (uint64_t)keys[i] - (uint64_t)base : 0;
        const bool act = i < n && (off0 >> 6) < (uint64_t)words;
        const uint64_t off = act ? off0 : 0;
        // Inactive lanes get word numbers no key has (distinct per lane).
        const uint64_t w = act ? (off >> 6) : ~(uint64_t)lane;
        uint64_t b = act ? (1ull << (off & 63)) : 0;
        for (int d = 1; d < 64; d <<= 1) {
            const uint64_t ow = __shfl_down(w, d, 64);
            const uint64_t ob = __shfl_down(b, d, 64);
            if (lane + d < 64 && ow == w) b |= ob;
        }
        const uint64_t pw = __shfl_up(w, 1, 64);
        if (act && (lane == 0 || pw != w)) atomicOr(&bits[w], b);
    }
}

// The probe form of a rank map: word w's bits and count side by side.
extern "C" __global__ void rf_rank_pack(const uint64_t* __restrict__ bits,
                                        const uint32_t* __restrict__ cum, int64_t words,
                                        unsigned long long* pack) {
    for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < words;
         w += (int64_t)gridDim.x * blockDim.x) {
        pack[2 * w] = bits[w];
        pack[2 * w + 1] = cum[w];
    }
}

// Sets *flag unless the keys are strictly ascending.
extern "C" __global__ void rf_not_ascending(const int64_t* __restrict__ keys, int64_t n,
                                            unsigned int* flag) {
    for (int64_t i = 1 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        if (keys[i] <= keys[i - 1]) atomicOr(flag, 1u);
}

// The first row of each distinct key at its rank in a built rank map:
// rows[rank] = min row holding the key (rows must hold INT64_MAX before).
extern "C" __global__ void rf_rank_rows(const int64_t* __restrict__ keys, int64_t n, int64_t base,
                                        const uint64_t* __restrict__ bits,
                                        const uint32_t* __restrict__ cum,
                                        unsigned long long* rows) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t off = (uint64_t)keys[i] - (uint64_t)base;
        const uint64_t wd = bits[off >> 6];
        const uint64_t r = (uint64_t)cum[off >> 6] + (uint64_t)__popcll(wd & ((1ull << (off & 63)) - 1));
        atomicMin(&rows[r], (unsigned long long)i);
    }
}

// Build rows in key order without a sort: each distinct key's rank is its
// position in the rank map (marked keys below it), so (key, row) goes to
// slot rank. Keys must be distinct (checked by the caller through the count
// of marked keys).
extern "C" __global__ void rf_rank_scatter(const int64_t* __restrict__ keys,
                                           const int64_t* __restrict__ rows, int64_t n,
                                           const int64_t* __restrict__ n_dev, int64_t base,
                                           const uint64_t* __restrict__ bits,
                                           const uint32_t* __restrict__ cum, int64_t* keys_out,
                                           int64_t* rows_out, int64_t words) {
    if (n_dev && *n_dev < n) n = *n_dev;  // as in rf_rank_bits
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = keys[i];
        const uint64_t off = (uint64_t)k - (uint64_t)base;
        if ((off >> 6) >= (uint64_t)words) continue;  // as in rf_rank_bits
        const uint64_t w = bits[off >> 6];
        const int64_t r = (int64_t)cum[off >> 6] + (int64_t)__popcll(w & ((1ull << (off & 63)) - 1));
        keys_out[r] = k;
        rows_out[r] = rows[i];
    }
}

// The number of marked keys: the last word's first rank plus its popcount.
extern "C" __global__ void rf_rank_total(const uint64_t* __restrict__ bits,
                                         const uint32_t* __restrict__ cum, int64_t words,
                                         uint64_t* out) {
    if (blockIdx.x == 0 && threadIdx.x == 0)
        *out = (uint64_t)cum[words - 1] + (uint64_t)__popcll(bits[words - 1]);
}

// A rank map's counts in one launch (one block of 1024 threads): cum[w] =
// marked keys in words before w, the packed probe form pack[2w] = bits[w],
// pack[2w + 1] = cum[w], and *total = all marked keys. A round covers
// 16 x 1024 words; thread t loads words r0 + j*1024 + t for every j before
// using any (coalesced loads, one memory wait per round), scans each j's
// 1024 counts within its wave, and wave 0 scans the 16 x 16 wave totals in
// word order: two barriers a round. (One word per thread per round, each
// round waiting on its loads, took 23 us for the 23K words of a 1.5M-key
// range; 16 consecutive words per thread, 50 us: uncoalesced, all on one CU.)
#define RF_RANK_PER 16
extern "C" __global__ __launch_bounds__(1024) void rf_rank_finish(
    const uint64_t* __restrict__ bits, int64_t words, uint32_t* __restrict__ cum,
    unsigned long long* __restrict__ pack, uint64_t* total, const int64_t* __restrict__ k_in,
    int64_t* k_out) {
    __shared__ uint32_t s_tot[RF_RANK_PER * 16];  // wave totals, then their prefixes
    __shared__ uint32_t s_round;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    uint32_t base = 0;
    for (int64_t r0 = 0; r0 < words; r0 += (int64_t)1024 * RF_RANK_PER) {
        uint64_t b[RF_RANK_PER];
#pragma unroll
        for (int j = 0; j < RF_RANK_PER; j++) {
            const int64_t w = r0 + (int64_t)j * 1024 + t;
            b[j] = w < words ? bits[w] : 0ull;
        }
        uint32_t inc[RF_RANK_PER];
#pragma unroll
        for (int j = 0; j < RF_RANK_PER; j++) {
            uint32_t x = (uint32_t)__popcll(b[j]);
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(x, o, 64);
                if (lane >= o) x += y;
            }
            inc[j] = x;
            if (lane == 63) s_tot[j * 16 + wv] = x;
        }
        __syncthreads();
        if (wv == 0) {
            // 256 wave totals in word order (j major), four per lane.
            const uint32_t v0 = s_tot[4 * lane], v1 = s_tot[4 * lane + 1],
                           v2 = s_tot[4 * lane + 2], v3 = s_tot[4 * lane + 3];
            const uint32_t sum = v0 + v1 + v2 + v3;
            uint32_t x = sum;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(x, o, 64);
                if (lane >= o) x += y;
            }
            const uint32_t ex = base + x - sum;
            s_tot[4 * lane] = ex;
            s_tot[4 * lane + 1] = ex + v0;
            s_tot[4 * lane + 2] = ex + v0 + v1;
            s_tot[4 * lane + 3] = ex + v0 + v1 + v2;
            if (lane == 63) s_round = x;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < RF_RANK_PER; j++) {
            const int64_t w = r0 + (int64_t)j * 1024 + t;
            const uint32_t below = s_tot[j * 16 + wv] + inc[j] - (uint32_t)__popcll(b[j]);
            if (w < words) {
                cum[w] = below;
                pack[2 * w] = b[j];
                pack[2 * w + 1] = below;
            }
        }
        base += s_round;
        __syncthreads();  // s_tot and s_round are rewritten next round
    }
    if (t == 0) {
        *total = base;
        if (k_in) *k_out = *k_in;  // the keys' count beside the marked keys' (pinned memory)
    }
}

extern "C" __global__ void rf_word_popc(const uint64_t* __restrict__ bits, int64_t words,
                                        uint32_t* cnt) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < words;
         i += (int64_t)gridDim.x * blockDim.x)
        cnt[i] = (uint32_t)__popcll(bits[i]);
}

// Concatenation of several runs' joined rows: blockIdx.y strides over the
// runs, blockIdx.x over a run's rows. Ranks get the run's first group added.
template <typename T>
__device__ __forceinline__ void rf_cat_copy(void* dst, const void* src, int64_t d, int64_t i) {
    ((T*)dst)[d] = ((const T*)src)[i];
}

extern "C" __global__ __launch_bounds__(256) void rf_gj_concat(const RfCatSeg* __restrict__ segs,
                                                               int64_t nseg, RfCatArgs A) {
    for (int64_t u = blockIdx.y; u < nseg; u += gridDim.y) {
        const RfCatSeg& S = segs[u];
        for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < S.n;
             i += (int64_t)gridDim.x * 256) {
            const int64_t d = S.dst + i;
            A.rank[d] = S.rank[i] + S.base;
            for (int j = 0; j < A.np; j++) {
                switch (A.width[j]) {
                    case 8: rf_cat_copy<uint64_t>(A.pay[j], S.pay[j], d, i); break;
                    case 4: rf_cat_copy<uint32_t>(A.pay[j], S.pay[j], d, i); break;
                    case 2: rf_cat_copy<uint16_t>(A.pay[j], S.pay[j], d, i); break;
                    default: rf_cat_copy<uint8_t>(A.pay[j], S.pay[j], d, i); break;
                }
            }
        }
    }
}

// ------------------------------------------- partitioned aggregation
//
// The default mode's group fold does not need the joined rows in probe-row
// order (only the deterministic mode does): the rows are counted per
// partition of 2^shift consecutive group ranks and moved into contiguous
// partitions, each block handling the staged rows of `waves_per_block` wave
// regions of the probe scan; the generated rf_gj_agg then folds a partition's
// groups in LDS. A block reserves its run of each partition with one atomic,
// so its rows of a partition land contiguously (runs of ~rows/nparts rows).

template <typename T>
__device__ __forceinline__ void rf_move(void* dst, const void* src, int64_t d, int64_t i) {
    ((T*)dst)[d] = ((const T*)src)[i];
}

#define RF_PART_BLOCK 1024
#define RF_PART_U 4

// The block's rows: the concatenation of its wave regions (at most 8);
// row k of the block is element rf_part_elem(k) of the staged columns. The
// region table lives in LDS (a private array indexed in a loop would be
// scratch memory).
struct RfPartRegions {
    int64_t base[8];
    int64_t end[8];  // cumulative row counts
    int n;
    int64_t total;
};

__device__ __forceinline__ void rf_part_regions(const RfPartArgs& P, RfPartRegions& R) {
    if (threadIdx.x == 0) {
        const int64_t w0 = (int64_t)blockIdx.x * P.waves_per_block;
        const int64_t w1 = w0 + P.waves_per_block < P.nwaves ? w0 + P.waves_per_block : P.nwaves;
        R.n = (int)(w1 - w0);
        int64_t tot = 0;
        for (int i = 0; i < R.n; i++) {
            R.base[i] = (w0 + i) * P.stage_rows;
            tot += P.wave_cnt[w0 + i];
            R.end[i] = tot;
        }
        R.total = tot;
    }
}

// Partitions that can hold rows: all nparts, or those below the count of
// groups when it is read from the device.
__device__ __forceinline__ int64_t rf_part_live(const RfPartArgs& P) {
    if (!P.groups_dev) return P.nparts;
    const int64_t n = (*P.groups_dev + ((int64_t)1 << P.shift) - 1) >> P.shift;
    return n < P.nparts ? n : P.nparts;
}

__device__ __forceinline__ int64_t rf_part_elem(const RfPartRegions& R, int64_t k) {
    int64_t prev = 0;
    for (int i = 0; i < R.n; i++) {
        if (k < R.end[i]) return R.base[i] + (k - prev);
        prev = R.end[i];
    }
    return -1;
}

// Rows per partition of each block: an LDS histogram per block, written to
// the block's column of P.boff (no global atomics: many blocks adding to the
// same few addresses serialise). Each thread issues RF_PART_U rank loads
// before using them.
extern "C" __global__ __launch_bounds__(RF_PART_BLOCK) void rf_part_hist(RfPartArgs P) {
    extern __shared__ uint32_t rf_ph[];
    const int t = threadIdx.x;
    const int64_t nl = rf_part_live(P);
    for (int64_t i = t; i < nl; i += RF_PART_BLOCK) rf_ph[i] = 0;
    __shared__ RfPartRegions R;
    rf_part_regions(P, R);
    __syncthreads();
    for (int64_t k0 = 0; k0 < R.total; k0 += (int64_t)RF_PART_BLOCK * RF_PART_U) {
        uint32_t r[RF_PART_U];
#pragma unroll
        for (int u = 0; u < RF_PART_U; u++) {
            const int64_t k = k0 + t + (int64_t)u * RF_PART_BLOCK;
            r[u] = k < R.total ? P.rank[rf_part_elem(R, k)] : 0xffffffffu;
        }
#pragma unroll
        for (int u = 0; u < RF_PART_U; u++)
            if (r[u] != 0xffffffffu) atomicAdd(&rf_ph[r[u] >> P.shift], 1u);
    }
    __syncthreads();
    for (int64_t i = t; i < nl; i += RF_PART_BLOCK)
        P.boff[i * P.nblocks + blockIdx.x] = rf_ph[i];
}

// Exclusive scan of one partition's block counts in place (block p), and
// the partition's row count.
extern "C" __global__ __launch_bounds__(256) void rf_part_rowscan(RfPartArgs P) {
    __shared__ uint64_t s_scan[4];
    if (blockIdx.x >= rf_part_live(P)) return;  // the whole block: an empty partition
    const int64_t nb = P.nblocks;
    uint32_t* row = P.boff + (int64_t)blockIdx.x * nb;
    const int64_t per = (nb + 255) / 256;
    const int64_t lo = (int64_t)threadIdx.x * per;
    const int64_t hi = lo + per < nb ? lo + per : nb;
    uint64_t mine = 0;
    for (int64_t i = lo; i < hi; i++) mine += row[i];
    uint64_t total;
    uint64_t off = rf_block_excl_scan<256>(mine, s_scan, total);
    for (int64_t i = lo; i < hi; i++) {
        const uint32_t c = row[i];
        row[i] = (uint32_t)off;
        off += c;
    }
    if (threadIdx.x == 0) P.ptot[blockIdx.x] = (uint32_t)total;
}

// Moves the block's rows into their partitions, CHUNK rows (RF_PART_CHUNK or
// twice that, P.chunk) at a time staged in LDS in partition order, so that
// consecutive threads write consecutive rows of a partition's run:
//  * the block's run of each partition is its scanned count in P.boff;
//  * per chunk: every row takes a slot in its partition (LDS counters), the
//    counters are scanned into the chunk's partition starts, the rows are
//    written to LDS in that order, and thread i writes LDS row i to its
//    run: run[p] + rows of p placed by earlier chunks + its place in p.
// Dynamic LDS: 4 * nparts counters and the chunk's rank + payload columns.
#define RF_PART_CHUNK 4096

#define RF_PART_PREG 2

__device__ __forceinline__ uint64_t rf_part_load(const void* src, int width, int64_t e) {
    switch (width) {
        case 8: return ((const uint64_t*)src)[e];
        case 4: return ((const uint32_t*)src)[e];
        case 2: return ((const uint16_t*)src)[e];
        default: return ((const uint8_t*)src)[e];
    }
}

template <typename T>
__device__ __forceinline__ void rf_part_from_lds(void* dst, const unsigned char* s, int64_t d, int i) {
    ((T*)dst)[d] = ((const T*)s)[i];
}

template <int CHUNK>
__device__ __forceinline__ void rf_part_scatter_body(const RfPartArgs& P) {
    extern __shared__ uint32_t rf_ps[];
    const int64_t np_ = P.nparts;
    uint32_t* cnt = rf_ps;              // block rows per partition, then rows placed so far
    uint32_t* run = rf_ps + np_;        // the block's run of each partition
    uint32_t* lc = rf_ps + 2 * np_;     // chunk rows per partition, then slots taken
    uint32_t* ls = rf_ps + 3 * np_;     // chunk start of each partition (LDS order)
    uint32_t* srank = rf_ps + 4 * np_;  // chunk ranks in partition order
    unsigned char* spay[8];
    {
        unsigned char* q = (unsigned char*)(srank + CHUNK);
        for (int j = 0; j < P.np; j++) {
            spay[j] = q;
            q += (size_t)CHUNK * P.width[j];
        }
    }
    __shared__ uint64_t s_scan[RF_PART_BLOCK / 64];
    const int t = threadIdx.x;
    __shared__ RfPartRegions R;
    rf_part_regions(P, R);
    // The block's run of each partition: the partition's start (a scan of
    // the partition totals, thread t taking a contiguous chunk of them) plus
    // the block's offset in it; block 0 publishes the starts for rf_gj_agg.
    const int64_t nl = rf_part_live(P);  // partitions that can hold rows
    const int64_t per = (nl + RF_PART_BLOCK - 1) / RF_PART_BLOCK;
    const int64_t lo = (int64_t)t * per;
    const int64_t hi = lo + per < nl ? lo + per : nl;
    uint64_t mine = 0, off, total;
    for (int64_t i = lo; i < hi; i++) mine += P.ptot[i];
    off = rf_block_excl_scan<RF_PART_BLOCK>(mine, s_scan, total);  // has barriers
    for (int64_t i = lo; i < hi; i++) {
        if (blockIdx.x == 0) P.pstart[i] = (uint32_t)off;
        run[i] = (uint32_t)off + P.boff[i * P.nblocks + blockIdx.x];
        cnt[i] = 0;
        off += P.ptot[i];
    }
    if (blockIdx.x == 0 && t == 0) P.pstart[nl] = (uint32_t)total;
    __syncthreads();
    constexpr int CU = CHUNK / RF_PART_BLOCK;
    for (int64_t k0 = 0; k0 < R.total; k0 += CHUNK) {
        for (int64_t i = t; i < nl; i += RF_PART_BLOCK) lc[i] = 0;
        __syncthreads();
        int64_t e[CU];
        uint32_t r[CU], slot[CU];
        // The first RF_PART_PREG payload columns are loaded with the ranks
        // (their latency overlaps); further ones when written to LDS.
        uint64_t pv[CU][RF_PART_PREG];
#pragma unroll
        for (int u = 0; u < CU; u++) {
            const int64_t k = k0 + t + (int64_t)u * RF_PART_BLOCK;
            e[u] = k < R.total ? rf_part_elem(R, k) : -1;
            r[u] = e[u] >= 0 ? P.rank[e[u]] : 0u;
#pragma unroll
            for (int j = 0; j < RF_PART_PREG; j++) {
                pv[u][j] = 0;
                if (j < P.np && e[u] >= 0) pv[u][j] = rf_part_load(P.pay[j], P.width[j], e[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < CU; u++)
            slot[u] = e[u] >= 0 ? atomicAdd(&lc[r[u] >> P.shift], 1u) : 0u;
        __syncthreads();
        mine = 0;
        for (int64_t i = lo; i < hi; i++) mine += lc[i];
        off = rf_block_excl_scan<RF_PART_BLOCK>(mine, s_scan, total);  // has barriers
        for (int64_t i = lo; i < hi; i++) {
            ls[i] = (uint32_t)off;
            off += lc[i];
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < CU; u++) {
            if (e[u] < 0) continue;
            const int i = (int)(ls[r[u] >> P.shift] + slot[u]);
            srank[i] = r[u];
            for (int j = 0; j < P.np; j++) {
                const uint64_t v = j < RF_PART_PREG ? pv[u][j < RF_PART_PREG ? j : 0]
                                                    : rf_part_load(P.pay[j], P.width[j], e[u]);
                switch (P.width[j]) {
                    case 8: ((uint64_t*)spay[j])[i] = v; break;
                    case 4: ((uint32_t*)spay[j])[i] = (uint32_t)v; break;
                    case 2: ((uint16_t*)spay[j])[i] = (uint16_t)v; break;
                    default: spay[j][i] = (uint8_t)v; break;
                }
            }
        }
        __syncthreads();
        const int n = (int)(R.total - k0 < CHUNK ? R.total - k0 : CHUNK);
        for (int i = t; i < n; i += RF_PART_BLOCK) {
            const uint32_t rr = srank[i];
            const uint32_t part = rr >> P.shift;
            const int64_t d = (int64_t)run[part] + cnt[part] + (uint32_t)(i - (int)ls[part]);
            P.prank[d] = rr;
            for (int j = 0; j < P.np; j++) {
                switch (P.width[j]) {
                    case 8: rf_part_from_lds<uint64_t>(P.ppay[j], spay[j], d, i); break;
                    case 4: rf_part_from_lds<uint32_t>(P.ppay[j], spay[j], d, i); break;
                    case 2: rf_part_from_lds<uint16_t>(P.ppay[j], spay[j], d, i); break;
                    default: rf_part_from_lds<uint8_t>(P.ppay[j], spay[j], d, i); break;
                }
            }
        }
        __syncthreads();
        for (int64_t i = t; i < nl; i += RF_PART_BLOCK) cnt[i] += lc[i];
    }
}

extern "C" __global__ __launch_bounds__(RF_PART_BLOCK) void rf_part_scatter(RfPartArgs P) {
    rf_part_scatter_body<RF_PART_CHUNK>(P);
}
extern "C" __global__ __launch_bounds__(RF_PART_BLOCK) void rf_part_scatter8k(RfPartArgs P) {
    rf_part_scatter_body<2 * RF_PART_CHUNK>(P);
}

// ------------------------------------------------------- host launchers

namespace rf {

static int grid_cap(int64_t n, int block, int cap) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

// Histogram entries a sort of n keys needs (digit counts + chunk sums).
int64_t aot_sort_hist_entries(int64_t n) {
    int64_t nb = (n + RF_SORT_TILE - 1) / RF_SORT_TILE;
    if (nb < 1) nb = 1;
    const int64_t m = 1024 * nb;  // room for 10-bit digits
    return m + (m + RF_SCAN_CHUNK - 1) / RF_SCAN_CHUNK;
}

// Stable LSD radix sort of n keys (4 or 8 bytes) by the low `bits` bits of
// (key - bias), moving up to 8 payload columns. keys[0]/pay[0][*] hold the
// input; the sorted result ends in buffer *which (0 or 1).
hipError_t aot_radix_sort(void* const keys[2], void* const pay[2][RF_SORT_MAXPAY],
                          const int32_t* pw, int np, int key_bytes, int64_t n, int bits,
                          uint64_t bias, uint32_t* hist, int* which, hipStream_t s) {
    *which = 0;
    if (n <= 1 || bits <= 0) return hipSuccess;
    if (np > RF_SORT_MAXPAY) return hipErrorInvalidValue;
    if (n >= (int64_t)1 << 32) return hipErrorInvalidValue;  // 32-bit positions
    RfSortPass P{};
    P.n = n;
    P.nblocks = (n + RF_SORT_TILE - 1) / RF_SORT_TILE;
    P.hist = hist;
    P.np = np;
    P.bias = bias;
    size_t elem = (size_t)key_bytes;
    for (int j = 0; j < np; j++) {
        P.pw[j] = pw[j];
        elem += (size_t)pw[j];
    }
    // Staged scatter when a tile's elements fit 48 KiB of LDS; then 10-bit
    // digits when they save a pass (17-20 bits: two passes instead of three).
    const size_t lds_bytes = (size_t)RF_SORT_TILE * elem;
    const bool staged = lds_bytes <= 49152;
    const bool wide = staged && (bits + 9) / 10 < (bits + 7) / 8;
    const int step = wide ? 10 : 8;
    int cur = 0;
    for (int shift = 0; shift < bits; shift += step) {
        P.kin = keys[cur];
        P.kout = keys[cur ^ 1];
        for (int j = 0; j < np; j++) {
            P.pin[j] = pay[cur][j];
            P.pout[j] = pay[cur ^ 1][j];
        }
        P.shift = shift;
        const dim3 g((unsigned)P.nblocks), b(RF_SORT_BLOCK);
        if (wide && key_bytes == 4)
            hipLaunchKernelGGL(rf_sort_hist10_u32, g, b, 0, s, P);
        else if (wide)
            hipLaunchKernelGGL(rf_sort_hist10_u64, g, b, 0, s, P);
        else if (key_bytes == 4)
            hipLaunchKernelGGL(rf_sort_hist_u32, g, b, 0, s, P);
        else
            hipLaunchKernelGGL(rf_sort_hist_u64, g, b, 0, s, P);
        const int64_t m = (int64_t)(1 << step) * P.nblocks;
        const int64_t nch = (m + RF_SCAN_CHUNK - 1) / RF_SCAN_CHUNK;
        uint32_t* part = hist + m;  // the caller sizes hist for m + chunks entries
        hipLaunchKernelGGL(rf_scan_chunks, dim3((unsigned)nch), dim3(256), 0, s, hist, m, part);
        hipLaunchKernelGGL(rf_sort_scan, dim3(1), dim3(1024), 0, s, part, nch);
        hipLaunchKernelGGL(rf_scan_apply, dim3((unsigned)nch), dim3(256), 0, s, hist, m, part);
        if (wide && key_bytes == 4)
            hipLaunchKernelGGL(rf_sort_scatter10_u32, g, b, lds_bytes, s, P);
        else if (wide)
            hipLaunchKernelGGL(rf_sort_scatter10_u64, g, b, lds_bytes, s, P);
        else if (staged && key_bytes == 4)
            hipLaunchKernelGGL(rf_sort_scatter_lds_u32, g, b, lds_bytes, s, P);
        else if (staged)
            hipLaunchKernelGGL(rf_sort_scatter_lds_u64, g, b, lds_bytes, s, P);
        else if (key_bytes == 4)
            hipLaunchKernelGGL(rf_sort_scatter_u32, g, b, 0, s, P);
        else
            hipLaunchKernelGGL(rf_sort_scatter_u64, g, b, 0, s, P);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        cur ^= 1;
    }
    *which = cur;
    return hipSuccess;
}

hipError_t aot_minmax_i64(const int64_t* keys, const int64_t* n_dev, int64_t n_max, void* out,
                          hipStream_t s) {
    hipLaunchKernelGGL(rf_minmax_i64, dim3(grid_cap(n_max, 256 * 16, 256)), dim3(256), 0, s, keys,
                       n_dev, (unsigned long long*)out);
    return hipGetLastError();
}

hipError_t aot_filter_build(const int64_t* keys, int64_t n, void* filter, uint64_t fmask,
                           hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(rf_filter_build, dim3(grid_cap(n, 256, 4096)), dim3(256), 0, s, keys, n,
                       (uint32_t*)filter, fmask);
    return hipGetLastError();
}

hipError_t aot_adjacent_dups(const int64_t* keys, int64_t n, void* flag, hipStream_t s) {
    if (n <= 1) return hipSuccess;
    hipLaunchKernelGGL(rf_adjacent_dups, dim3(grid_cap(n, 256, 4096)), dim3(256), 0, s, keys, n,
                       (unsigned int*)flag);
    return hipGetLastError();
}

hipError_t aot_seg_bounds(const uint32_t* rank, int64_t n, int64_t* lo, int64_t* hi,
                          hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(rf_seg_bounds, dim3(grid_cap(n, 256, 16384)), dim3(256), 0, s, rank, n,
                       lo, hi);
    return hipGetLastError();
}

// Keys and rows in key order from unsorted distinct keys and a built rank
// map, and the number of marked keys into *total.
hipError_t aot_rank_scatter(const int64_t* keys, const int64_t* rows, int64_t n, int64_t base,
                            const void* bits, int64_t words, const void* cum, int64_t* keys_out,
                            int64_t* rows_out, void* total, hipStream_t s) {
    if (n <= 0 || words <= 0) return hipSuccess;
    hipLaunchKernelGGL(rf_rank_scatter, dim3(grid_cap(n, 256, 4096)), dim3(256), 0, s, keys, rows, n,
                       (const int64_t*)nullptr, base, (const uint64_t*)bits, (const uint32_t*)cum,
                       keys_out, rows_out, words);
    hipLaunchKernelGGL(rf_rank_total, dim3(1), dim3(64), 0, s, (const uint64_t*)bits,
                       (const uint32_t*)cum, words, (uint64_t*)total);
    return hipGetLastError();
}

// Entries the rank map's cumulative counts need for `words` words (the
// counts and the scan's chunk sums).
int64_t aot_rank_cum_entries(int64_t words) {
    return words + (words + RF_SCAN_CHUNK - 1) / RF_SCAN_CHUNK + 1;
}

// bits: `words` zeroed words; cum: aot_rank_cum_entries(words) entries.
hipError_t aot_rank_build(const int64_t* keys, int64_t n, int64_t base, void* bits, int64_t words,
                          void* cum, hipStream_t s) {
    if (n <= 0 || words <= 0) return hipSuccess;
    hipLaunchKernelGGL(rf_rank_bits, dim3(grid_cap(n, 256, 4096)), dim3(256), 0, s, keys, n,
                       (const int64_t*)nullptr, base, (unsigned long long*)bits, words);
    hipLaunchKernelGGL(rf_word_popc, dim3(grid_cap(words, 256, 4096)), dim3(256), 0, s,
                       (const uint64_t*)bits, words, (uint32_t*)cum);
    uint32_t* c = (uint32_t*)cum;
    const int64_t nch = (words + RF_SCAN_CHUNK - 1) / RF_SCAN_CHUNK;
    uint32_t* part = c + words;
    hipLaunchKernelGGL(rf_scan_chunks, dim3((unsigned)nch), dim3(256), 0, s, c, words, part);
    hipLaunchKernelGGL(rf_sort_scan, dim3(1), dim3(1024), 0, s, part, nch);
    hipLaunchKernelGGL(rf_scan_apply, dim3((unsigned)nch), dim3(256), 0, s, c, words, part);
    return hipGetLastError();
}

// The rank map of n keys in words zeroed words, counts and packed form in
// two launches (marking, then rf_rank_finish): for maps of at most
// kRankFusedWords words, where one block scans them quickly. *total gets
// the number of marked keys. With n_dev set, the keys are the first
// min(*n_dev, n) (a count a kernel before wrote: no host synchronisation),
// and with k_out set, rf_rank_finish copies *k_in there.
hipError_t aot_rank_build_small(const int64_t* keys, int64_t n, const int64_t* n_dev, int64_t base,
                                void* bits, int64_t words, void* cum, void* pack, void* total,
                                const int64_t* k_in, int64_t* k_out, hipStream_t s) {
    if (words <= 0) return hipSuccess;
    if (n > 0)
        hipLaunchKernelGGL(rf_rank_bits, dim3(grid_cap(n, 256, 4096)), dim3(256), 0, s, keys, n,
                           n_dev, base, (unsigned long long*)bits, words);
    hipLaunchKernelGGL(rf_rank_finish, dim3(1), dim3(1024), 0, s, (const uint64_t*)bits, words,
                       (uint32_t*)cum, (unsigned long long*)pack, (uint64_t*)total,
                       k_out ? k_in : (const int64_t*)nullptr, k_out);
    return hipGetLastError();
}

// Places (key, row) at each key's rank (rf_rank_scatter alone); n_dev as in
// aot_rank_build_small.
hipError_t aot_rank_place(const int64_t* keys, const int64_t* rows, int64_t n, const int64_t* n_dev,
                          int64_t base, const void* bits, int64_t words, const void* cum,
                          int64_t* keys_out, int64_t* rows_out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(rf_rank_scatter, dim3(grid_cap(n, 256, 4096)), dim3(256), 0, s, keys, rows, n,
                       n_dev, base, (const uint64_t*)bits, (const uint32_t*)cum, keys_out,
                       rows_out, words);
    return hipGetLastError();
}

hipError_t aot_rank_pack(const void* bits, const void* cum, int64_t words, void* pack,
                         hipStream_t s) {
    if (words <= 0) return hipSuccess;
    hipLaunchKernelGGL(rf_rank_pack, dim3(grid_cap(words, 256, 4096)), dim3(256), 0, s,
                       (const uint64_t*)bits, (const uint32_t*)cum, words,
                       (unsigned long long*)pack);
    return hipGetLastError();
}

// Merge-path splits of a union merge's tiles: out[b] = the A keys among the
// first min(b * tile, na + nb) keys of the merge of ascending A and B.
extern "C" __global__ void rf_merge_splits(const int64_t* __restrict__ A, int64_t na,
                                           const int64_t* __restrict__ B, int64_t nb,
                                           uint64_t oa, uint64_t ob, int64_t tile, int64_t n,
                                           int64_t* __restrict__ out) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const int64_t m = na + nb;
    out[b] = rf_merge_split(A, na, B, nb, b * tile < m ? b * tile : m, oa, ob);
}

hipError_t aot_merge_splits(const int64_t* a, int64_t na, const int64_t* b, int64_t nb,
                            int64_t off_a, int64_t off_b, int64_t tile, int64_t n, int64_t* out,
                            hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(rf_merge_splits, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, na,
                       b, nb, (uint64_t)off_a, (uint64_t)off_b, tile, n, out);
    return hipGetLastError();
}

// Entries the exclusive scan of m uint32 counts needs (the counts and the
// scan's chunk sums).
int64_t aot_scan_entries(int64_t m) { return m + (m + RF_SCAN_CHUNK - 1) / RF_SCAN_CHUNK; }

// Exclusive scan in place of h[0, m) (h holds aot_scan_entries(m) entries).
hipError_t aot_scan_u32(void* h, int64_t m, hipStream_t s) {
    if (m <= 0) return hipSuccess;
    uint32_t* c = (uint32_t*)h;
    const int64_t nch = (m + RF_SCAN_CHUNK - 1) / RF_SCAN_CHUNK;
    uint32_t* part = c + m;
    hipLaunchKernelGGL(rf_scan_chunks, dim3((unsigned)nch), dim3(256), 0, s, c, m, part);
    hipLaunchKernelGGL(rf_sort_scan, dim3(1), dim3(1024), 0, s, part, nch);
    hipLaunchKernelGGL(rf_scan_apply, dim3((unsigned)nch), dim3(256), 0, s, c, m, part);
    return hipGetLastError();
}

hipError_t aot_not_ascending(const int64_t* keys, int64_t n, void* flag, hipStream_t s) {
    if (n <= 1) return hipSuccess;
    hipLaunchKernelGGL(rf_not_ascending, dim3(grid_cap(n, 256, 4096)), dim3(256), 0, s, keys, n,
                       (unsigned int*)flag);
    return hipGetLastError();
}

hipError_t aot_rank_rows(const int64_t* keys, int64_t n, int64_t base, const void* bits,
                         const void* cum, int64_t* rows, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(rf_rank_rows, dim3(grid_cap(n, 256, 4096)), dim3(256), 0, s, keys, n, base,
                       (const uint64_t*)bits, (const uint32_t*)cum, (unsigned long long*)rows);
    return hipGetLastError();
}

// Counts and moves a probe scan's staged joined rows into partitions of
// 2^shift ranks (P.hist and P.cursor zeroed before; at most 8192 partitions).
// Entries (uint32) of the block counts, partition totals and starts for
// nparts partitions of the joined rows of nwaves wave regions: boff is at
// 0, ptot at nparts * nblocks, pstart after it.
int64_t aot_part_entries(int64_t nparts, int64_t nwaves, int64_t waves_per_block) {
    const int64_t nb = (nwaves + waves_per_block - 1) / waves_per_block;
    return nparts * nb + 2 * nparts + 1;
}

hipError_t aot_partition(const RfPartArgs& P0, hipStream_t s) {
    RfPartArgs P = P0;
    P.nblocks = (P.nwaves + P.waves_per_block - 1) / P.waves_per_block;
    P.ptot = P.boff + P.nparts * P.nblocks;
    P.pstart = P.ptot + P.nparts;
    if (P.nparts <= 0 || P.nparts > 4096 || P.nwaves <= 0) return hipErrorInvalidValue;
    if (P.waves_per_block < 1 || P.waves_per_block > 8) return hipErrorInvalidValue;
    const int64_t nb = (P.nwaves + P.waves_per_block - 1) / P.waves_per_block;
    P.nblocks = nb;
    if (!P.counted)
        hipLaunchKernelGGL(rf_part_hist, dim3((unsigned)nb), dim3(RF_PART_BLOCK),
                           (size_t)P.nparts * 4, s, P);
    // Each block's offset in each partition and the partition totals.
    hipLaunchKernelGGL(rf_part_rowscan, dim3((unsigned)P.nparts), dim3(256), 0, s, P);
    size_t row = 4;
    for (int j = 0; j < P.np; j++) row += (size_t)P.width[j];
    const int chunk = P.chunk == 2 * RF_PART_CHUNK ? 2 * RF_PART_CHUNK : RF_PART_CHUNK;
    const void* kern = chunk == RF_PART_CHUNK ? (const void*)rf_part_scatter
                                              : (const void*)rf_part_scatter8k;
    const size_t lds = (size_t)P.nparts * 16 + (size_t)chunk * row;
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    if (chunk == RF_PART_CHUNK)
        hipLaunchKernelGGL(rf_part_scatter, dim3((unsigned)nb), dim3(RF_PART_BLOCK), lds, s, P);
    else
        hipLaunchKernelGGL(rf_part_scatter8k, dim3((unsigned)nb), dim3(RF_PART_BLOCK), lds, s, P);
    return hipGetLastError();
}

// segs: device array of nseg runs; max_rows: the largest run.
hipError_t aot_gj_concat(const RfCatSeg* segs, int32_t nseg, int64_t max_rows,
                         const RfCatArgs& A, hipStream_t s) {
    if (nseg <= 0 || max_rows <= 0) return hipSuccess;
    const int gx = grid_cap(max_rows, 1024, 256);
    const int gy = nseg < 65535 ? nseg : 65535;
    hipLaunchKernelGGL(rf_gj_concat, dim3(gx, gy), dim3(256), 0, s, segs, (int64_t)nseg, A);
    return hipGetLastError();
}

}  // namespace rf
