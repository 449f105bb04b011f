// Hand-written gfx950 scan skeleton for lowered reffine queries.
//
// This header is compiled by hiprtc together with a generated `struct Q`
// that supplies the per-row code of one query (column loads, predicate,
// per-field contributions, selection outputs). Everything that is not
// row-specific lives here: the static tile schedule, 128-bit coalesced
// column loads, the fixed-order wave/block/grid reductions, the ordered
// stream compaction (per-block staging in tile order + a gather in tile
// order) and the hash probe of reffine's build_index
// (include/reffine/arrow/table.h:107-138).
//
// Row semantics follow the reference's generated loop
// (src/pass/loopgen.cpp:206-269): a row contributes to state field f iff
// its predicate holds, and the field folds with +, *, min, max, && or ||.
//
// Fixed reduction order (the deterministic mode; independent of grid size
// and of which block runs which tile):
//   * a tile is SUB x BLOCK x ITEMS consecutive rows; in sub-step s thread
//     t owns rows tile0 + (s*BLOCK + t)*ITEMS + [0, ITEMS);
//   * each thread folds its rows in ascending order, starting from the
//     fold's identity;
//   * a wave combines its 64 lanes with xor-butterflies 1,2,4,...,32
//     (lane value v_i <- v_i (+) v_{i^o}); the block folds its waves in
//     order w0 (+) w1 (+) ...; the result is tile partial[tile];
//   * the completing block: thread t folds partial[t], partial[t+BLOCK],
//     ... in order, then the same wave/block combine gives the total.
#pragma once

#include "rf_abi.h"

// ------------------------------------------------------------ loads

template <typename T, int N>
struct RfVec {
    T v[N];
};

// One row block of N elements starting at an element index that is a
// multiple of N; N*sizeof(T) is a multiple of 4 here. Loads in 16-byte
// pieces (global_load_dwordx4) where the width allows.
template <typename T, int N>
__device__ __forceinline__ RfVec<T, N> rf_ldv(const T* __restrict__ p) {
    RfVec<T, N> r;
    constexpr int BYTES = N * (int)sizeof(T);
    if constexpr (BYTES % 16 == 0) {
        typedef unsigned int u4 __attribute__((ext_vector_type(4)));
        const u4* q = reinterpret_cast<const u4*>(p);
        u4 tmp[BYTES / 16];
#pragma unroll
        for (int i = 0; i < BYTES / 16; i++) tmp[i] = __builtin_nontemporal_load(q + i);
        __builtin_memcpy(&r, tmp, BYTES);
    } else if constexpr (BYTES % 8 == 0) {
        typedef unsigned int u2 __attribute__((ext_vector_type(2)));
        const u2* q = reinterpret_cast<const u2*>(p);
        u2 tmp[BYTES / 8];
#pragma unroll
        for (int i = 0; i < BYTES / 8; i++) tmp[i] = __builtin_nontemporal_load(q + i);
        __builtin_memcpy(&r, tmp, BYTES);
    } else {
#pragma unroll
        for (int i = 0; i < N; i++) r.v[i] = p[i];
    }
    return r;
}

template <typename T, int N>
__device__ __forceinline__ RfVec<T, N> rf_ldv_tail(const T* __restrict__ p,
                                                   int64_t r, int64_t end) {
    RfVec<T, N> x;
#pragma unroll
    for (int i = 0; i < N; i++) x.v[i] = (r + i < end) ? p[r + i] : T(0);
    return x;
}

// Validity bits of rows [r, r+N) from a uint16-word bitmap (LSB first, the
// reference's get_null_bit, include/reffine/vinstr/vinstr.h:70-74). N
// divides 16 and r is a multiple of N, so one word holds them all.
template <int N>
__device__ __forceinline__ uint32_t rf_valid_bits(const uint16_t* __restrict__ bm,
                                                  int64_t r) {
    uint32_t w = bm[r >> 4];
    return (w >> (uint32_t)(r & 15)) & ((1u << N) - 1u);
}

__device__ __forceinline__ bool rf_valid_bit(const uint16_t* __restrict__ bm,
                                             int64_t i) {
    return (bm[i >> 4] >> (uint32_t)(i & 15)) & 1u;
}

// ------------------------------------------------------------ probes

__device__ __forceinline__ uint64_t rf_hash64(uint64_t k) {
    // MurmurHash3 fmix64 finaliser.
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

__device__ __forceinline__ int64_t rf_probe(const RfProbeDev& h, int64_t key) {
    if (key == RF_EMPTY_KEY) return h.min_key_row;
    uint64_t s = rf_hash64((uint64_t)key) & h.mask;
    for (;;) {
        int64_t k = h.keys[s];
        if (k == key) return h.rows[s];
        if (k == RF_EMPTY_KEY) return -1;
        s = (s + 1) & h.mask;
    }
}

// ------------------------------------------------------------ helpers

template <typename T>
__device__ __forceinline__ T rf_max(T a, T b) { return (a < b) ? b : a; }  // std::max
template <typename T>
__device__ __forceinline__ T rf_min(T a, T b) { return (b < a) ? b : a; }  // std::min
template <typename T>
__device__ __forceinline__ T rf_abs(T a) { return (a < T(0)) ? -a : a; }

template <typename T>
__device__ __forceinline__ T rf_shfl_xor(T v, int o) {
    if constexpr (sizeof(T) == 8) {
        int64_t x = __builtin_bit_cast(int64_t, v);
        x = __shfl_xor(x, o, 64);
        return __builtin_bit_cast(T, x);
    } else {
        int x = (int)v;
        x = __shfl_xor(x, o, 64);
        return (T)x;
    }
}

__device__ __forceinline__ uint32_t rf_lane() { return __lane_id(); }

// ---------------------------------------------------- block reductions

template <class Q>
__device__ __forceinline__ void rf_wave_reduce(typename Q::Acc& a) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) Q::acc_shfl_xor(a, o);
}

// Reduces `a` over the block in the fixed order; valid in thread 0.
template <class Q>
__device__ __forceinline__ void rf_block_reduce(typename Q::Acc& a,
                                                typename Q::Acc* s_waves) {
    rf_wave_reduce<Q>(a);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) s_waves[w] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        a = s_waves[0];
#pragma unroll
        for (int i = 1; i < Q::BLOCK / 64; i++) Q::acc_combine(a, s_waves[i]);
    }
    __syncthreads();
}

// Block-wide exclusive scan of 64-bit values (packed 16-bit counters);
// returns the thread's exclusive prefix and writes the block total.
template <int BLOCK>
__device__ __forceinline__ uint64_t rf_block_excl_scan(uint64_t x, uint64_t* s_tmp,
                                                       uint64_t& total) {
    const uint32_t lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    uint64_t inc = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint64_t y = __shfl_up(inc, o, 64);
        if (lane >= (uint32_t)o) inc += y;
    }
    if (lane == 63) s_tmp[w] = inc;
    __syncthreads();
    uint64_t wave_base = 0;
    total = 0;
#pragma unroll
    for (int i = 0; i < BLOCK / 64; i++) {
        uint64_t t = s_tmp[i];
        if (i < w) wave_base += t;
        total += t;
    }
    __syncthreads();
    return wave_base + inc - x;
}

// ------------------------------------------------------------ the kernel

// Thread t's exclusive share of a chunked sum: thread t owns entries
// [t*per, (t+1)*per) of `v`; returns the prefix of the entries before its
// chunk and writes the total. Used by the completing block to scan tile
// counts in tile order.
template <int BLOCK>
__device__ __forceinline__ uint64_t rf_chunk_prefix(const int64_t* v, int64_t n, int64_t per,
                                                    uint64_t* s_tmp, uint64_t& total) {
    uint64_t mine = 0;
    const int64_t lo = (int64_t)threadIdx.x * per;
    for (int64_t i = lo; i < lo + per && i < n; i++) mine += (uint64_t)v[i];
    return rf_block_excl_scan<BLOCK>(mine, s_tmp, total);
}

// Scan kernel. Tiles are assigned statically (tile = blockIdx.x + k *
// gridDim.x), so no tile waits on another and no atomics are taken per
// tile. A selection is compacted per block: block b appends the selected
// rows of its tiles, in tile order, to its own staging region
// [b * stage_rows, (b + 1) * stage_rows) and records each tile's staging
// offset and count; the completing block turns the counts into output
// offsets (in tile order) and rf_gather moves the rows into place.
template <class Q>
__device__ __forceinline__ void rf_scan(const RfScanArgs& L) {
    typedef typename Q::Acc Acc;
    constexpr int BLOCK = Q::BLOCK, ITEMS = Q::ITEMS, SUB = Q::SUB;
    constexpr int64_t TILE = (int64_t)SUB * BLOCK * ITEMS;
    __shared__ Acc s_waves[BLOCK / 64];
    __shared__ uint64_t s_scan[BLOCK / 64];
    __shared__ uint32_t s_last;

    Acc run;
    Q::acc_zero(run);
    int64_t cursor = (int64_t)blockIdx.x * L.stage_rows;  // block-uniform
    for (int64_t tile = blockIdx.x; tile < L.ntiles; tile += gridDim.x) {
        const int64_t row0 = L.start + tile * TILE;
        Acc a;
        Q::acc_zero(a);
        uint32_t bits[SUB];
        if (row0 >= L.lo && row0 + TILE <= L.end) {
            // Interior tile: all SUB sub-steps' loads are issued before any
            // row is evaluated.
            Q::tile_full(a, L, row0, bits);
        } else {
#pragma unroll
            for (int s = 0; s < SUB; s++) {
                const int64_t r = row0 + ((int64_t)s * BLOCK + threadIdx.x) * ITEMS;
                if (r >= L.lo && r + ITEMS <= L.end)
                    bits[s] = Q::template rows<true>(a, L, r);
                else if (r + ITEMS > L.lo && r < L.end)
                    bits[s] = Q::template rows<false>(a, L, r);
                else
                    bits[s] = 0;
            }
        }
        if constexpr (Q::NACC > 0) {
            if constexpr (Q::DET) {
                rf_block_reduce<Q>(a, s_waves);
                if (threadIdx.x == 0) reinterpret_cast<Acc*>(L.partials)[tile] = a;
            } else {
                Q::acc_combine(run, a);
            }
        }
        if constexpr (Q::HAS_SEL) {
            // Rows of a tile in ascending order are (sub-step, thread, item):
            // one 16-bit counter per sub-step, packed, scanned across the
            // block in a single pass.
            uint64_t packed = 0;
#pragma unroll
            for (int s = 0; s < SUB; s++)
                packed |= (uint64_t)__builtin_popcount(bits[s]) << (16 * s);
            uint64_t total;
            const uint64_t excl = rf_block_excl_scan<BLOCK>(packed, s_scan, total);
            uint64_t sbase = 0;
#pragma unroll
            for (int s = 0; s < SUB; s++) {
                uint32_t b = bits[s];
                int64_t pos = cursor + (int64_t)(sbase + ((excl >> (16 * s)) & 0xffffull));
                const int64_t r = row0 + ((int64_t)s * BLOCK + threadIdx.x) * ITEMS;
                while (b) {
                    const int k = __builtin_ctz(b);
                    b &= b - 1;
                    Q::emit(L, r + k, pos);
                    pos++;
                }
                sbase += (total >> (16 * s)) & 0xffffull;
            }
            if (threadIdx.x == 0) {
                L.tile_src[tile] = cursor;
                L.tile_cnt[tile] = (int64_t)sbase;
            }
            cursor += (int64_t)sbase;
        }
        // Closes the tile. Keeps every thread-0 region of one iteration apart
        // from the next iteration's, so the compiler cannot thread them
        // together around the barriers.
        __syncthreads();
    }

    if constexpr (Q::NACC > 0 && !Q::DET) {
        rf_block_reduce<Q>(run, s_waves);
        if (threadIdx.x == 0) reinterpret_cast<Acc*>(L.partials)[blockIdx.x] = run;
    }

    // Completion: the last block to finish combines the partials, turns the
    // tile counts into output offsets and resets the counter for the next
    // launch.
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t prev = atomicAdd(&L.counters[1], 1u);
        s_last = (prev == gridDim.x - 1) ? 1u : 0u;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    if constexpr (Q::NACC > 0) {
        const int64_t nparts = Q::DET ? L.ntiles : (int64_t)gridDim.x;
        const Acc* parts = reinterpret_cast<const Acc*>(L.partials);
        Acc a;
        Q::acc_zero(a);
        for (int64_t i = threadIdx.x; i < nparts; i += BLOCK) Q::acc_combine(a, parts[i]);
        rf_block_reduce<Q>(a, s_waves);
        if (threadIdx.x == 0) *reinterpret_cast<Acc*>(L.result) = a;
    }
    if constexpr (Q::HAS_SEL) {
        const int64_t per = (L.ntiles + BLOCK - 1) / BLOCK;
        uint64_t total;
        uint64_t off = rf_chunk_prefix<BLOCK>(L.tile_cnt, L.ntiles, per, s_scan, total);
        const int64_t lo = (int64_t)threadIdx.x * per;
        for (int64_t i = lo; i < lo + per && i < L.ntiles; i++) {
            L.tile_dst[i] = (int64_t)off;
            off += (uint64_t)L.tile_cnt[i];
        }
        if (threadIdx.x == 0) *L.out_count = (int64_t)total;
    }
    if (threadIdx.x == 0) {
        L.counters[1] = 0;
        __threadfence();
    }
}

// Moves each tile's staged rows to its output offset (one tile per block
// iteration, threads over the rows).
template <class Q>
__device__ __forceinline__ void rf_gather(const RfScanArgs& L) {
    for (int64_t tile = blockIdx.x; tile < L.ntiles; tile += gridDim.x) {
        const int64_t src = L.tile_src[tile], dst = L.tile_dst[tile], cnt = L.tile_cnt[tile];
        for (int64_t i = threadIdx.x; i < cnt; i += blockDim.x) Q::move(L, src + i, dst + i);
    }
}
