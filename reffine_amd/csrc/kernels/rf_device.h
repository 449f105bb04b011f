// Hand-written gfx950 scan skeleton for lowered reffine queries.
//
// This header is compiled by hiprtc together with a generated `struct Q`
// that supplies the per-row code of one query (column loads, predicate,
// per-field contributions, selection outputs). Everything that is not
// row-specific lives here: the dynamic tile scheduler, 128-bit coalesced
// column loads, the fixed-order wave/block/grid reductions, the ordered
// stream compaction (ballot/popcount + decoupled look-back) and the hash
// probe of reffine's build_index (include/reffine/arrow/table.h:107-138).
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

// ------------------------------------------- decoupled look-back status
// Status word: [63:62] flag (1 aggregate, 2 inclusive prefix), [61:40]
// launch epoch, [39:0] value.
#define RF_ST_AGG 1ull
#define RF_ST_PRE 2ull
#define RF_ST_VAL_MASK ((1ull << 40) - 1)
#define RF_ST_EPOCH_MASK ((1ull << 22) - 1)

__device__ __forceinline__ uint64_t rf_st_make(uint64_t flag, uint64_t epoch,
                                               uint64_t val) {
    return (flag << 62) | ((epoch & RF_ST_EPOCH_MASK) << 40) | (val & RF_ST_VAL_MASK);
}

// A status word carries flag, epoch and value in one 64-bit atomic and no
// other data is published through it, so relaxed device-scope atomics
// suffice (no L2 writeback / invalidate per tile).
__device__ __forceinline__ void rf_st_store(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t rf_st_load(const uint64_t* p) {
    return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

// Upper bound on polls of one predecessor's status word. Predecessors are
// always resident (tiles are handed out in order by an atomic ticket), so
// the bound is only a safety net: past it the wave stops waiting, raises
// the error word and the host reports the run as failed.
#define RF_SPIN_LIMIT (1u << 24)

// Wave 0 computes the exclusive prefix of tile `tile` given its aggregate;
// publishes aggregate then inclusive prefix. Returns the exclusive prefix
// (valid in all lanes of wave 0).
__device__ __forceinline__ uint64_t rf_lookback(uint64_t* status, int64_t tile,
                                                uint64_t agg, uint64_t epoch,
                                                uint32_t* err, uint32_t spin_limit) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t ep = epoch & RF_ST_EPOCH_MASK;
    if (tile == 0) {
        if (lane == 0) rf_st_store(status, rf_st_make(RF_ST_PRE, ep, agg));
        return 0;
    }
    if (lane == 0) rf_st_store(status + tile, rf_st_make(RF_ST_AGG, ep, agg));
    uint64_t excl = 0;
    int64_t base = tile - 1;
    for (;;) {
        const int64_t t = base - (int64_t)lane;
        uint64_t st = 0;
        uint64_t flag = 0;
        if (t >= 0) {
            for (uint32_t spins = 0;; spins++) {
                st = rf_st_load(status + t);
                flag = st >> 62;
                if (flag != 0 && ((st >> 40) & RF_ST_EPOCH_MASK) == ep) break;
                if (spins >= spin_limit) {
                    atomicOr(err, 1u);
                    flag = RF_ST_PRE;
                    st = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        } else {
            flag = RF_ST_PRE;  // before tile 0: prefix 0
            st = 0;
        }
        const uint64_t val = (t >= 0) ? (st & RF_ST_VAL_MASK) : 0;
        const uint64_t pre_mask = __ballot(flag == RF_ST_PRE);
        // Closest predecessor holding an inclusive prefix.
        const int stop = pre_mask ? __builtin_ctzll(pre_mask) : 64;
        uint64_t contrib = ((int)lane <= stop) ? val : 0;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) contrib += __shfl_xor(contrib, o, 64);
        excl += contrib;
        if (pre_mask) break;
        base -= 64;
    }
    if (lane == 0) rf_st_store(status + tile, rf_st_make(RF_ST_PRE, ep, excl + agg));
    return excl;
}

// ------------------------------------------------------------ the kernel

template <class Q>
__device__ __forceinline__ void rf_scan(const RfScanArgs& L) {
    typedef typename Q::Acc Acc;
    constexpr int BLOCK = Q::BLOCK, ITEMS = Q::ITEMS, SUB = Q::SUB;
    constexpr int64_t TILE = (int64_t)SUB * BLOCK * ITEMS;
    __shared__ Acc s_waves[BLOCK / 64];
    __shared__ uint64_t s_scan[BLOCK / 64];
    __shared__ int64_t s_tile[2];
    __shared__ uint64_t s_prefix;
    __shared__ uint32_t s_last;

    Acc run;
    Q::acc_zero(run);
    // Loop structure: every iteration ends with a barrier and the tile index
    // is made scalar, so the loop exit is a uniform scalar branch. Without
    // the closing barrier the compiler may merge the thread-0 region at the
    // end of one iteration with the thread-0 region at the start of the
    // next, sending lane 0 and lanes 1..63 of wave 0 around different paths
    // through the barriers.
    // The next tile's ticket is fetched while the current tile is processed
    // (its latency hides behind the tile's loads). Tickets alternate between
    // two slots: iteration i reads slot i&1 and writes slot (i+1)&1, and the
    // closing barrier orders every read before the slot is written again.
    if (threadIdx.x == 0) s_tile[0] = (int64_t)atomicAdd(&L.counters[0], 1u);
    __syncthreads();
    for (int par = 0;; par ^= 1) {
        const int64_t tile_v = s_tile[par];
        const int64_t tile =
            (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(tile_v >> 32)) << 32) |
                      (uint32_t)__builtin_amdgcn_readfirstlane((int)tile_v));
        if (tile >= L.ntiles) break;
        uint32_t next_ticket = 0;
        if (threadIdx.x == 0) next_ticket = atomicAdd(&L.counters[0], 1u);
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
            uint64_t packed = 0;
#pragma unroll
            for (int s = 0; s < SUB; s++)
                packed |= (uint64_t)__builtin_popcount(bits[s]) << (16 * s);
            uint64_t total;
            const uint64_t excl = rf_block_excl_scan<BLOCK>(packed, s_scan, total);
            uint64_t tile_count = 0;
#pragma unroll
            for (int s = 0; s < SUB; s++) tile_count += (total >> (16 * s)) & 0xffffull;
            if (threadIdx.x < 64) {
                const uint64_t pre = rf_lookback(L.tile_status, tile, tile_count, L.epoch,
                                                 &L.counters[2], L.spin_limit);
                if (threadIdx.x == 0) s_prefix = pre;
            }
            __syncthreads();
            const uint64_t prefix = s_prefix;
            uint64_t sbase = 0;
#pragma unroll
            for (int s = 0; s < SUB; s++) {
                uint32_t b = bits[s];
                uint64_t pos = prefix + sbase + ((excl >> (16 * s)) & 0xffffull);
                const int64_t r = row0 + ((int64_t)s * BLOCK + threadIdx.x) * ITEMS;
                while (b) {
                    const int k = __builtin_ctz(b);
                    b &= b - 1;
                    Q::emit(L, r + k, (int64_t)pos);
                    pos++;
                }
                sbase += (total >> (16 * s)) & 0xffffull;
            }
        }
        // Closes the tile: s_prefix and s_scan may be rewritten next; s_tile
        // takes the prefetched ticket (read by all only after the barrier).
        if (threadIdx.x == 0) s_tile[par ^ 1] = (int64_t)next_ticket;
        __syncthreads();
    }

    if constexpr (Q::NACC > 0 && !Q::DET) {
        rf_block_reduce<Q>(run, s_waves);
        if (threadIdx.x == 0) reinterpret_cast<Acc*>(L.partials)[blockIdx.x] = run;
    }

    // Completion: the last block to finish combines the partials, writes the
    // selection count and resets the counters for the next launch.
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
    if (threadIdx.x == 0) {
        if constexpr (Q::HAS_SEL) {
            uint64_t cnt = 0;
            if (L.ntiles > 0) cnt = rf_st_load(L.tile_status + (L.ntiles - 1)) & RF_ST_VAL_MASK;
            *L.out_count = (int64_t)cnt;
        }
        L.counters[0] = 0;
        L.counters[1] = 0;
        __threadfence();
    }
}
