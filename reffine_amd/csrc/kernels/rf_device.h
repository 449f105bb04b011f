// Hand-written gfx950 scan skeleton for lowered reffine queries.
//
// This header is compiled by hiprtc together with a generated `struct Q`
// that supplies the per-row code of one query (column loads, predicate,
// per-field contributions, selection outputs). Everything that is not
// row-specific lives here: the wave-level tile schedule, 128-bit coalesced
// column loads, the fixed-order wave/grid reductions, wavefront
// ballot/popcount stream compaction, and the hash probe of reffine's
// build_index (include/reffine/arrow/table.h:107-138).
//
// Row semantics follow the reference's generated loop
// (src/pass/loopgen.cpp:206-269): a row contributes to state field f iff
// its predicate holds, and the field folds with +, *, min, max, && or ||.
//
// Work unit: a wave-tile of SUB x 64 x ITEMS consecutive rows; in sub-step
// s lane l owns rows tile0 + (s*64 + l)*ITEMS + [0, ITEMS). Each wave owns a
// contiguous range of wave-tiles, so the steady state has no block barrier.
//
// Fixed reduction order (the deterministic mode; independent of the grid
// and of which wave runs which tile):
//   * each lane folds its rows of a tile in ascending order, starting from
//     the fold's identity;
//   * the wave combines its 64 lanes with xor-butterflies 1,2,4,...,32
//     (lane value v_i <- v_i (+) v_{i^o}); lane 0's value is partial[tile];
//   * the completing block: thread t folds partial[t], partial[t+BLOCK],
//     ... in order; each wave combines its lanes by the same butterflies and
//     the block folds its waves in order w0 (+) w1 (+) ...
#pragma once

#include "rf_abi.h"

// ------------------------------------------------------------ loads

// Every buffer the kernels touch is global memory. Pointers read from
// memory (a batch's per-unit arguments) lose that address space, which would
// make their accesses flat ones; these casts restore it.
#define RF_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const RF_GLOBAL T* rf_g(const T* p) {
    return (const RF_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ RF_GLOBAL T* rf_gw(T* p) {
    return (RF_GLOBAL T*)p;
}

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
        const RF_GLOBAL u4* q = reinterpret_cast<const RF_GLOBAL u4*>(rf_g(p));
        u4 tmp[BYTES / 16];
#pragma unroll
        for (int i = 0; i < BYTES / 16; i++) tmp[i] = __builtin_nontemporal_load(q + i);
        __builtin_memcpy(&r, tmp, BYTES);
    } else if constexpr (BYTES % 8 == 0) {
        typedef unsigned int u2 __attribute__((ext_vector_type(2)));
        const RF_GLOBAL u2* q = reinterpret_cast<const RF_GLOBAL u2*>(rf_g(p));
        u2 tmp[BYTES / 8];
#pragma unroll
        for (int i = 0; i < BYTES / 8; i++) tmp[i] = __builtin_nontemporal_load(q + i);
        __builtin_memcpy(&r, tmp, BYTES);
    } else {
#pragma unroll
        for (int i = 0; i < N; i++) r.v[i] = rf_g(p)[i];
    }
    return r;
}

template <typename T, int N>
__device__ __forceinline__ RfVec<T, N> rf_ldv_tail(const T* __restrict__ p,
                                                   int64_t r, int64_t end) {
    RfVec<T, N> x;
#pragma unroll
    for (int i = 0; i < N; i++) x.v[i] = (r + i < end) ? rf_g(p)[r + i] : T(0);
    return x;
}

// Validity bits of rows [r, r+N) from a uint16-word bitmap (LSB first, the
// reference's get_null_bit, include/reffine/vinstr/vinstr.h:70-74). N
// divides 16 and r is a multiple of N, so one word holds them all.
template <int N>
__device__ __forceinline__ uint32_t rf_valid_bits(const uint16_t* __restrict__ bm,
                                                  int64_t r) {
    uint32_t w = rf_g(bm)[r >> 4];
    return (w >> (uint32_t)(r & 15)) & ((1u << N) - 1u);
}

__device__ __forceinline__ bool rf_valid_bit(const uint16_t* __restrict__ bm,
                                             int64_t i) {
    return (rf_g(bm)[i >> 4] >> (uint32_t)(i & 15)) & 1u;
}

// ------------------------------------------------------------ merges

// Merge-path split: how many of A's keys are among the first d keys of the
// merge of ascending A (na keys, each + oa) and B (nb keys, each + ob) that
// takes A's key first on a tie. Reads A and B only inside [0, na) and
// [0, nb) for 0 <= d <= na + nb. The offsets add in wrapping arithmetic.
__device__ __forceinline__ int64_t rf_merge_split(const int64_t* __restrict__ A, int64_t na,
                                                  const int64_t* __restrict__ B, int64_t nb,
                                                  int64_t d, uint64_t oa, uint64_t ob) {
    int64_t lo = d > nb ? d - nb : 0, hi = d < na ? d : na;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)((uint64_t)rf_g(A)[mid] + oa) <= (int64_t)((uint64_t)rf_g(B)[d - mid - 1] + ob))
            lo = mid + 1;
        else hi = mid;
    }
    return lo;
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

// The bits and count of a rank map's word w in one 16-byte load.
__device__ __forceinline__ void rf_rank_word(const RfProbeDev& h, uint64_t w, uint64_t& bits,
                                             uint32_t& below) {
    typedef unsigned long long u2 __attribute__((ext_vector_type(2)));
    const u2 v = rf_g(reinterpret_cast<const u2*>(h.dense_pack))[w];
    bits = v.x;
    below = (uint32_t)v.y;
}

__device__ __forceinline__ int64_t rf_probe(const RfProbeDev& h, int64_t key) {
    if (h.dense_pack) {
        const uint64_t off = (uint64_t)key - (uint64_t)h.dense_base;
        if (off >= h.dense_len) return -1;
        uint64_t wd;
        uint32_t below;
        rf_rank_word(h, off >> 6, wd, below);
        const uint64_t bit = 1ull << (off & 63);
        if (!(wd & bit)) return -1;
        const int64_t r = (int64_t)below + (int64_t)__popcll(wd & (bit - 1));
        return h.dense_rows ? rf_g(h.dense_rows)[r] : r;
    }
    if (h.dense_bits) {
        // Rank of the key among the marked keys (a small, cache-resident
        // structure); the offset wraps to >= dense_len for keys below the base.
        const uint64_t off = (uint64_t)key - (uint64_t)h.dense_base;
        if (off >= h.dense_len) return -1;
        const uint64_t wd = rf_g(h.dense_bits)[off >> 6];
        const uint64_t bit = 1ull << (off & 63);
        if (!(wd & bit)) return -1;
        const int64_t r = (int64_t)rf_g(h.dense_cum)[off >> 6] + (int64_t)__popcll(wd & (bit - 1));
        return h.dense_rows ? rf_g(h.dense_rows)[r] : r;
    }
    if (key == RF_EMPTY_KEY) return h.min_key_row;
    const uint64_t hk = rf_hash64((uint64_t)key);
    if (h.filter) {
        // Keys absent from the filter are absent from the table: one read
        // of a small (cache-resident) bitmap instead of a slot read.
        const uint64_t fb = (hk >> 32) & h.fmask;
        if (!((rf_g(h.filter)[fb >> 5] >> (fb & 31)) & 1u)) return -1;
    }
    uint64_t s = hk & h.mask;
    for (;;) {
        int64_t k = rf_g(h.keys)[s];
        if (k == key) return rf_g(h.rows)[s];
        if (k == RF_EMPTY_KEY) return -1;
        s = (s + 1) & h.mask;
    }
}

// A rank-map probe of the scan's row code: the row of `key`, or -1 when the
// row's guard fails or the key is absent. The map's packed form (word bits
// and count in one 16-byte load) is read for every key in the range; the
// probe must hold a rank map with its packed form (every rank map the
// library builds has one).
__device__ __forceinline__ int64_t rf_probe_dense(const RfProbeDev& h, int64_t key, bool guard) {
    if (!guard) return -1;
    const uint64_t off = (uint64_t)key - (uint64_t)h.dense_base;
    if (off >= h.dense_len) return -1;
    uint64_t wd;
    uint32_t below;
    rf_rank_word(h, off >> 6, wd, below);
    const uint64_t bit = 1ull << (off & 63);
    if (!(wd & bit)) return -1;
    return (int64_t)below + (int64_t)__popcll(wd & (bit - 1));
}

// The packed rank-map probe without branches, for a map with ascending keys
// (no row table) and at least one key: a key outside the map reads word 0
// and is then rejected. Straight-line code lets a fold's consecutive rows
// issue their probes together.
__device__ __forceinline__ int64_t rf_probe_pack(const RfProbeDev& h, int64_t key, bool guard) {
    const uint64_t off = (uint64_t)key - (uint64_t)h.dense_base;
    const bool in = guard && off < h.dense_len;
    const uint64_t o = in ? off : 0;
    uint64_t wd;
    uint32_t below;
    rf_rank_word(h, o >> 6, wd, below);
    const uint64_t bit = 1ull << (o & 63);
    const int64_t r = (int64_t)below + (int64_t)__popcll(wd & (bit - 1));
    return (in && (wd & bit) != 0) ? r : (int64_t)-1;
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
    } else if constexpr (sizeof(T) == 4) {
        // Bits, not a value conversion: a float partial moves unchanged.
        return __builtin_bit_cast(T, __shfl_xor(__builtin_bit_cast(int, v), o, 64));
    } else {
        return (T)__shfl_xor((int)v, o, 64);
    }
}

// The value of the lane that DPP control CTRL selects (inside a 16-lane row).
template <int CTRL, typename T>
__device__ __forceinline__ T rf_dpp(T v) {
    if constexpr (sizeof(T) == 8) {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        const uint32_t lo =
            (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, true);
        const uint32_t hi =
            (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, true);
        return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
    } else if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(
            T, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
    } else {
        return (T)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, true);
    }
}

// A butterfly level that must pair lanes exactly as xor XOR does: the DPP
// control for XOR == 0 (the quad permutations are xor 1 and 2), else the
// xor shuffle (the mirrors pair the same lanes only when every lane of a
// quad holds the same bits, which a non-commutative combine breaks).
template <int CTRL, int XOR, typename T>
__device__ __forceinline__ T rf_dpp_x(T v) {
    if constexpr (XOR == 0) return rf_dpp<CTRL>(v);
    else return rf_shfl_xor(v, XOR);
}

// Lane k's value (k wave-uniform), for every lane.
template <typename T>
__device__ __forceinline__ T rf_readlane(T v, int k) {
    if constexpr (sizeof(T) == 8) {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, k);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), k);
        return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
    } else if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), k));
    } else {
        return (T)__builtin_amdgcn_readlane((int)v, k);
    }
}

__device__ __forceinline__ uint32_t rf_lane() { return __lane_id(); }

// ---------------------------------------------------- block reductions

template <class Q>
__device__ __forceinline__ void rf_wave_reduce(typename Q::Acc& a) {
    // The xor butterfly o = 1, 2, 4, ..., 32. The first four levels move
    // data with DPP inside 16-lane rows (no LDS round trip): xor 1 and 2 as
    // quad permutations; then, with every lane of a quad holding the quad's
    // value, the half-row mirror pairs quads as xor 4 does, and the row
    // mirror pairs half-rows as xor 8 does -- for a commutative combine the
    // lanes hold exactly the butterfly's values (float min/max fields, not
    // commutative with NaNs, take xor shuffles there). xor 16 and 32 cross
    // rows.
    Q::template acc_dpp<0xB1, 0>(a);   // quad_perm(1,0,3,2): xor 1
    Q::template acc_dpp<0x4E, 0>(a);   // quad_perm(2,3,0,1): xor 2
    Q::template acc_dpp<0x141, 4>(a);  // row_half_mirror (xor 4)
    Q::template acc_dpp<0x140, 8>(a);  // row_mirror (xor 8)
    Q::acc_shfl_xor(a, 16);
    Q::acc_shfl_xor(a, 32);
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

__device__ __forceinline__ int64_t rf_scalar64(int64_t v) {
    return (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(v >> 32)) << 32) |
                     (uint32_t)__builtin_amdgcn_readfirstlane((int)v));
}

// Count of set bits of `mask` in lanes below this lane.
__device__ __forceinline__ uint32_t rf_lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// A wave-tile (the fold unit of the deterministic order: SUB sub-steps of
// 64 lanes x ITEMS rows) is evaluated in load groups of LSUB sub-steps (the
// registers one load holds); in sub-step s of a group starting at row0,
// lane l owns rows row0 + (s*64 + l)*ITEMS + [0, ITEMS). A tile of NG
// groups therefore gives each lane the same rows, in the same ascending
// order, as one group of SUB sub-steps would.

// Evaluates one load group starting at row0: folds its rows into `a`,
// writes the selection bits of each sub-step and the per-row values the
// selection's emit reuses. Interior groups use the registers loaded into
// `T`; edge groups load with bounds checks.
template <class Q>
__device__ __forceinline__ void rf_tile_rows(const RfScanArgs& L, const typename Q::Tile& T,
                                             int64_t row0, bool interior, uint32_t lane,
                                             typename Q::Acc& a, uint32_t* bits,
                                             typename Q::Carry& cy) {
    constexpr int ITEMS = Q::ITEMS, LSUB = Q::LSUB;
    if (interior) {
        Q::compute(a, T, L, row0, lane, bits, cy);
    } else {
#pragma unroll
        for (int s = 0; s < LSUB; s++) {
            const int64_t r = row0 + ((int64_t)s * 64 + lane) * ITEMS;
            if (r >= L.lo && r + ITEMS <= L.end)
                bits[s] = Q::template rows<true>(a, L, r, cy, s);
            else if (r + ITEMS > L.lo && r < L.end)
                bits[s] = Q::template rows<false>(a, L, r, cy, s);
            else
                bits[s] = 0;
        }
    }
}

// Appends a load group's qualifying rows, in row order, to the staging
// columns at positions base + cursor, ...; advances the wave-uniform
// `cursor`. Rows of a sub-step in ascending order are (lane, item): the
// position of (lane, k) is the number of qualifying rows in lower lanes (one
// ballot per item) plus those of items < k in the lane.
template <class Q>
__device__ __forceinline__ void rf_tile_emit(const RfScanArgs& L, int64_t row0, uint32_t lane,
                                             const uint32_t* bits, const typename Q::Carry& cy,
                                             int64_t base, int64_t& cursor) {
    constexpr int ITEMS = Q::ITEMS, LSUB = Q::LSUB;
#pragma unroll
    for (int s = 0; s < LSUB; s++) {
        const uint32_t b = bits[s];
        uint32_t below = 0, total = 0;
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const uint64_t m = __ballot((b >> k) & 1u);
            below += rf_lanes_below(m);
            total += (uint32_t)__builtin_popcountll(m);
        }
        if (b) {
            const int64_t r = row0 + ((int64_t)s * 64 + lane) * ITEMS;
            int64_t pos = base + cursor + below;
#pragma unroll
            for (int k = 0; k < ITEMS; k++) {
                if ((b >> k) & 1u) {
                    Q::emit(L, r + k, pos, cy, s, k);
                    pos++;
                }
            }
        }
        cursor += total;
    }
}

// Partitions the fused histogram of a group join's probe scan counts.
__device__ __forceinline__ int64_t rf_scan_parts(const RfScanArgs& L) {
    if (!L.part_groups_dev) return L.part_nparts;
    const int64_t n = (*L.part_groups_dev + ((int64_t)1 << L.part_shift) - 1) >> L.part_shift;
    return n < L.part_nparts ? n : L.part_nparts;
}

// rf_tile_emit through the wave's LDS buffer: the load group's qualifying
// rows are written to LDS at their place among the group's rows (the same
// row order), then the wave writes them to the staging columns 64 rows per
// store. LDS operations of one wave complete in issue order, so lanes read
// what other lanes of the wave wrote before; the wave barriers keep the
// compiler from moving the buffer's reads and writes across the hand-over.
template <class Q>
__device__ __forceinline__ void rf_tile_emit_lds(const RfScanArgs& L, int64_t row0, uint32_t lane,
                                                 const uint32_t* bits, const typename Q::Carry& cy,
                                                 int64_t base, int64_t& cursor,
                                                 typename Q::EmitBuf& buf, uint32_t* hist) {
    constexpr int ITEMS = Q::ITEMS, LSUB = Q::LSUB;
    uint32_t n = 0;  // the group's qualifying rows so far (wave-uniform)
#pragma unroll
    for (int s = 0; s < LSUB; s++) {
        const uint32_t b = bits[s];
        uint32_t below = 0, total = 0;
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const uint64_t m = __ballot((b >> k) & 1u);
            below += rf_lanes_below(m);
            total += (uint32_t)__builtin_popcountll(m);
        }
        if (b) {
            const int64_t r = row0 + ((int64_t)s * 64 + lane) * ITEMS;
            int i = (int)(n + below);
#pragma unroll
            for (int k = 0; k < ITEMS; k++) {
                if ((b >> k) & 1u) {
                    Q::emit_lds(buf, i, r + k, cy, s, k);
                    i++;
                }
            }
        }
        n += total;
    }
    if (n == 0) return;
    __builtin_amdgcn_wave_barrier();
    for (uint32_t i = lane; i < n; i += 64) {
        Q::flush(L, buf, (int)i, base + cursor + i);
        if (L.part_boff) atomicAdd(&hist[(uint32_t)Q::part_key(buf, (int)i) >> L.part_shift], 1u);
    }
    __builtin_amdgcn_wave_barrier();
    cursor += n;
}

// A load group's joined rows folded into the block's LDS slots (a group
// join's probe pre-aggregating in LDS): no staging.
template <class Q>
__device__ __forceinline__ void rf_tile_agg(const RfScanArgs& L, const uint32_t* bits,
                                            const typename Q::Carry& cy, unsigned long long* acc) {
    constexpr int ITEMS = Q::ITEMS, LSUB = Q::LSUB;
#pragma unroll
    for (int s = 0; s < LSUB; s++) {
        const uint32_t b = bits[s];
        if (!b) continue;
#pragma unroll
        for (int k = 0; k < ITEMS; k++)
            if ((b >> k) & 1u) Q::agg_row(L, acc, L.agg_gpp, cy, s, k);
    }
}

// Scan kernel. Wave w (of W = gridDim.x * BLOCK/64) owns wave-tiles
// [w*per, min((w+1)*per, ntiles)) with per = ceil(ntiles / W), processed as
// a sequence of load groups.
//   * reductions fold into registers (per wave-tile partials in the
//     deterministic mode);
//   * a selection is compacted per wave with ballot/popcount: the wave
//     appends its qualifying rows, in row order, to its staging region
//     [w*stage_rows, (w+1)*stage_rows) and records its count; as waves own
//     contiguous ascending ranges, the output is the concatenation of the
//     waves' staged rows in wave order. The completing block scans the W
//     counts into offsets and rf_gather moves each wave's rows into place.
template <class Q>
__device__ __forceinline__ void rf_scan(const RfScanArgs& L) {
    typedef typename Q::Acc Acc;
    constexpr int BLOCK = Q::BLOCK, ITEMS = Q::ITEMS, LSUB = Q::LSUB;
    constexpr int NG = Q::SUB / Q::LSUB;  // load groups per wave-tile
    constexpr int WAVES = BLOCK / 64;
    constexpr int64_t GROUP = (int64_t)LSUB * 64 * ITEMS;
    __shared__ Acc s_waves[WAVES];
    __shared__ uint64_t s_scan[WAVES];
    __shared__ uint32_t s_last;
    __shared__ typename Q::EmitBuf s_emit[Q::LDS_EMIT ? WAVES : 1];  // per-wave emit buffers
    // Dynamic LDS: the LDS_AGG slots (NSLOT x agg_gpp), or the fused
    // partition histogram of an LDS_EMIT probe scan (part_nparts counters).
    extern __shared__ unsigned long long rf_agg_lds[];
    uint32_t* const rf_hist = reinterpret_cast<uint32_t*>(rf_agg_lds);
    if constexpr (Q::LDS_AGG) {
        for (int64_t i = threadIdx.x; i < (int64_t)Q::NSLOT * L.agg_gpp; i += BLOCK)
            rf_agg_lds[i] = Q::agg_ident((int)(i / L.agg_gpp));
        __syncthreads();
    }
    if constexpr (Q::LDS_EMIT) {
        if (L.part_boff) {  // uniform: a kernel argument
            const int64_t np = rf_scan_parts(L);
            for (int64_t i = threadIdx.x; i < np; i += BLOCK) rf_hist[i] = 0u;
            __syncthreads();
        }
    }

    const uint32_t lane = __lane_id();
    const int64_t gw = (int64_t)blockIdx.x * WAVES + (threadIdx.x >> 6);
    const int64_t nw = (int64_t)gridDim.x * WAVES;
    const int64_t per = (L.ntiles + nw - 1) / nw;
    const int64_t t_begin = rf_scalar64(gw * per);
    const int64_t t_end = rf_scalar64(t_begin + per < L.ntiles ? t_begin + per : L.ntiles);
    const int64_t g_begin = t_begin * NG, g_end = t_end * NG;  // the wave's load groups

    Acc run;
    Q::acc_zero(run);
    Acc a_tile;  // deterministic mode: the current tile's fold
    Q::acc_zero(a_tile);
    int64_t cursor = 0;  // rows this wave has staged (wave-uniform)

    // One load group: rows from the loaded registers `T` (interior groups)
    // or with bounds checks (edge groups), then its fold and compaction. In
    // the deterministic mode each tile folds from the identity and its last
    // group writes the tile's partial; otherwise rows fold straight into
    // the wave's running state.
    auto process = [&](const typename Q::Tile& T, int64_t grp, bool interior) {
        const int64_t row0 = L.start + grp * GROUP;
        Acc& a = Q::DET ? a_tile : run;
        uint32_t bits[LSUB];
        typename Q::Carry cy;  // per-row values the selection's emit reuses
        rf_tile_rows<Q>(L, T, row0, interior, lane, a, bits, cy);
        if constexpr (Q::LDS_AGG) rf_tile_agg<Q>(L, bits, cy, rf_agg_lds);
        if constexpr (Q::HAS_SEL) {
            if constexpr (Q::LDS_EMIT)
                rf_tile_emit_lds<Q>(L, row0, lane, bits, cy, gw * L.stage_rows, cursor,
                                    s_emit[threadIdx.x >> 6], rf_hist);
            else
                rf_tile_emit<Q>(L, row0, lane, bits, cy, gw * L.stage_rows, cursor);
        }
        if constexpr (Q::NACC > 0 && Q::DET) {
            if (grp % NG == NG - 1) {
                Acc t = a_tile;
                rf_wave_reduce<Q>(t);
                if (lane == 0) rf_gw(reinterpret_cast<Acc*>(L.partials))[grp / NG] = t;
                Q::acc_zero(a_tile);
            }
        }
    };
    auto interior = [&](int64_t grp) {
        const int64_t row0 = L.start + grp * GROUP;
        return row0 >= L.lo && row0 + GROUP <= L.end;
    };

    if constexpr (Q::PIPE) {
        // Software pipeline over the wave's load groups, two register sets:
        // the next group's loads are in flight while the current group is
        // evaluated.
        typename Q::Tile TA, TB;
        if (g_begin < g_end && interior(g_begin)) Q::load(TA, L, L.start + g_begin * GROUP, lane);
        bool in0 = g_begin < g_end && interior(g_begin);
        for (int64_t grp = g_begin; grp < g_end; grp++) {
            const bool in1 = grp + 1 < g_end && interior(grp + 1);
            if (in1) Q::load(TB, L, L.start + (grp + 1) * GROUP, lane);
            process(TA, grp, in0);
            // One copy of the row code per iteration (the instruction cache
            // holds it); the next group's registers move into place.
            TA = TB;
            in0 = in1;
        }
    } else {
        for (int64_t grp = g_begin; grp < g_end; grp++) {
            typename Q::Tile T;
            const bool in = interior(grp);
            if (in) Q::load(T, L, L.start + grp * GROUP, lane);
            process(T, grp, in);
        }
    }

    if constexpr (Q::HAS_SEL) {
        if (lane == 0) L.wave_cnt[gw] = cursor;
    }
    if constexpr (Q::LDS_EMIT) {
        if (L.part_boff) {
            // The block's column of the partition count matrix.
            __syncthreads();
            const int64_t np = rf_scan_parts(L);
            for (int64_t i = threadIdx.x; i < np; i += BLOCK)
                L.part_boff[i * (int64_t)gridDim.x + blockIdx.x] = rf_hist[i];
        }
    }
    if constexpr (Q::LDS_AGG) {
        // The block's slots, for rf_gj_lds_done to combine across blocks.
        __syncthreads();
        unsigned long long* dst = L.agg_part + (int64_t)blockIdx.x * Q::NSLOT * L.agg_gpp;
        for (int64_t i = threadIdx.x; i < (int64_t)Q::NSLOT * L.agg_gpp; i += BLOCK)
            dst[i] = rf_agg_lds[i];
    }
    if constexpr (Q::NACC > 0 && !Q::DET) {
        rf_wave_reduce<Q>(run);
        if (lane == 0) reinterpret_cast<Acc*>(L.partials)[gw] = run;
    }

    // Completion: the last block to finish combines the partials, turns the
    // wave counts into output offsets and resets the counter for the next
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
        const int64_t nparts = Q::DET ? L.ntiles : nw;
        const Acc* parts = reinterpret_cast<const Acc*>(L.partials);
        Acc a;
        Q::acc_zero(a);
        for (int64_t i = threadIdx.x; i < nparts; i += BLOCK) Q::acc_combine(a, parts[i]);
        rf_block_reduce<Q>(a, s_waves);
        if (threadIdx.x == 0) *reinterpret_cast<Acc*>(L.result) = a;
    }
    if constexpr (Q::HAS_SEL) {
        // Exclusive scan of the nw wave counts: thread t takes a contiguous
        // chunk of `pt` of them.
        const int64_t pt = (nw + BLOCK - 1) / BLOCK;
        const int64_t lo = (int64_t)threadIdx.x * pt;
        uint64_t mine = 0;
        for (int64_t i = lo; i < lo + pt && i < nw; i++) mine += (uint64_t)L.wave_cnt[i];
        uint64_t total;
        uint64_t off = rf_block_excl_scan<BLOCK>(mine, s_scan, total);
        for (int64_t i = lo; i < lo + pt && i < nw; i++) {
            L.wave_dst[i] = (int64_t)off;
            off += (uint64_t)L.wave_cnt[i];
        }
        if (threadIdx.x == 0) *L.out_count = (int64_t)total;
    }
    if (threadIdx.x == 0) {
        L.counters[1] = 0;
        __threadfence();
    }
}

// Moves each wave's staged rows to its output offset. Rows at offsets
// beyond out_cap are not placed (the host sized the output from an
// estimate and places them again into a larger table).
template <class Q>
__device__ __forceinline__ void rf_gather(const RfScanArgs& L) {
    for (int64_t w = blockIdx.x; w < L.nwaves; w += gridDim.x) {
        const int64_t src = w * L.stage_rows, dst = rf_g(L.wave_dst)[w], cnt = rf_g(L.wave_cnt)[w];
        const int64_t room = L.out_cap - dst;
        const int64_t n = cnt < room ? cnt : room;
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x) Q::move(L, src + i, dst + i);
    }
}

// ------------------------------------------------------ batched execution
//
// Many work units (one query over one record batch each, an RfScanArgs per
// unit in device memory) run in one persistent launch over the list of
// (unit, wave-tile) items; each wave takes a contiguous range of the list.
// Every item is processed
// as the deterministic scan processes a tile: its fold from the identity,
// wave-combined into partials[tile]; its qualifying rows staged at
// tile * TILE with the count in wave_cnt[tile] (so stage_rows = TILE and
// nwaves = ntiles for a unit). rf_batch_done then combines each unit's
// partials in the deterministic scan's fixed order and turns the tile
// counts into offsets; rf_gather places the rows. A unit therefore gets
// the same bits as the deterministic scan of the same query alone.

template <class Q>
__device__ __forceinline__ void rf_batch_scan(const RfBatchArgs& B) {
    typedef typename Q::Acc Acc;
    constexpr int ITEMS = Q::ITEMS, SUB = Q::SUB, LSUB = Q::LSUB;
    constexpr int NG = SUB / LSUB;
    constexpr int WAVES = Q::BLOCK / 64;
    constexpr int64_t GROUP = (int64_t)LSUB * 64 * ITEMS;
    constexpr int64_t TILE = (int64_t)SUB * 64 * ITEMS;
    const uint32_t lane = __lane_id();
    const int64_t nunits = B.hdr->nunits, nitems = B.hdr->nitems;
    const RfScanArgs* work = (const RfScanArgs*)(B.work + B.hdr->unit_base);
    const int64_t* item_start = B.item_start + B.hdr->start_base;
    // Items are the units' wave-tiles, concatenated; wave w takes the
    // contiguous range [w*per, (w+1)*per) (one counter shared by every wave
    // would serialise on its address).
    const int64_t gw = (int64_t)blockIdx.x * WAVES + (threadIdx.x >> 6);
    const int64_t nw = (int64_t)gridDim.x * WAVES;
    const int64_t per = (nitems + nw - 1) / nw;
    int64_t item = rf_scalar64(gw * per);
    const int64_t item_end = rf_scalar64(item + per < nitems ? item + per : nitems);
    if (item >= item_end) return;
    // Unit of the first item: the last u with item_start[u] <= item.
    int64_t u = 0, hi = nunits - 1;
    while (u < hi) {
        const int64_t mid = (u + hi + 1) >> 1;
        if (item_start[mid] <= item) u = mid;
        else hi = mid - 1;
    }
    u = rf_scalar64(u);
    for (; item < item_end; item++) {
        while (item_start[u + 1] <= item) u++;
        const RfScanArgs& L = work[u];
        const int64_t tile = item - item_start[u];
        Acc a;
        Q::acc_zero(a);
        int64_t cursor = 0;
#pragma unroll 1
        for (int g = 0; g < NG; g++) {
            const int64_t row0 = L.start + tile * TILE + g * GROUP;
            const bool interior = row0 >= L.lo && row0 + GROUP <= L.end;
            typename Q::Tile T;
            if (interior) Q::load(T, L, row0, lane);
            uint32_t bits[LSUB];
            typename Q::Carry cy;
            rf_tile_rows<Q>(L, T, row0, interior, lane, a, bits, cy);
            if constexpr (Q::HAS_SEL) rf_tile_emit<Q>(L, row0, lane, bits, cy, tile * TILE, cursor);
        }
        if constexpr (Q::NACC > 0) {
            rf_wave_reduce<Q>(a);
            if (lane == 0) rf_gw(reinterpret_cast<Acc*>(L.partials))[tile] = a;
        }
        if constexpr (Q::HAS_SEL) {
            if (lane == 0) rf_gw(L.wave_cnt)[tile] = cursor;
        }
    }
}

// One block per unit: the deterministic scan's completion for that unit.
template <class Q>
__device__ __forceinline__ void rf_batch_done(const RfBatchArgs& B) {
    typedef typename Q::Acc Acc;
    constexpr int BLOCK = Q::BLOCK;
    __shared__ Acc s_waves[BLOCK / 64];
    __shared__ uint64_t s_scan[BLOCK / 64];
    const int64_t u = blockIdx.x;
    if (u >= B.hdr->nunits) return;
    const RfScanArgs& L = *(const RfScanArgs*)(B.work + B.hdr->unit_base + u);
    if constexpr (Q::NACC > 0) {
        const RF_GLOBAL Acc* parts = rf_g(reinterpret_cast<const Acc*>(L.partials));
        Acc a;
        Q::acc_zero(a);
        for (int64_t i = threadIdx.x; i < L.ntiles; i += BLOCK) {
            const Acc p = parts[i];
            Q::acc_combine(a, p);
        }
        rf_block_reduce<Q>(a, s_waves);
        if (threadIdx.x == 0) *rf_gw(reinterpret_cast<Acc*>(L.result)) = a;
    }
    if constexpr (Q::HAS_SEL) {
        const int64_t n = L.nwaves;
        const int64_t pt = (n + BLOCK - 1) / BLOCK;
        const int64_t lo = (int64_t)threadIdx.x * pt;
        const RF_GLOBAL int64_t* cnt = rf_g(L.wave_cnt);
        RF_GLOBAL int64_t* dst = rf_gw(L.wave_dst);
        uint64_t mine = 0;
        for (int64_t i = lo; i < lo + pt && i < n; i++) mine += (uint64_t)cnt[i];
        uint64_t total;
        uint64_t off = rf_block_excl_scan<BLOCK>(mine, s_scan, total);
        for (int64_t i = lo; i < lo + pt && i < n; i++) {
            dst[i] = (int64_t)off;
            off += (uint64_t)cnt[i];
        }
        if (threadIdx.x == 0) *rf_gw(L.out_count) = (int64_t)total;
    }
}

// Selection placement of every unit: a persistent launch over the same
// (unit, wave-tile) item list as the scan, one wave per tile's staged rows,
// each wave a contiguous range of items.
template <class Q>
__device__ __forceinline__ void rf_batch_gather(const RfBatchArgs& B) {
    constexpr int64_t TILE = (int64_t)Q::SUB * 64 * Q::ITEMS;
    const uint32_t lane = __lane_id();
    const int64_t nunits = B.hdr->nunits, nitems = B.hdr->nitems;
    const RfScanArgs* work = (const RfScanArgs*)(B.work + B.hdr->unit_base);
    const int64_t* item_start = B.item_start + B.hdr->start_base;
    const int64_t gw = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x / 64);
    const int64_t per = (nitems + nw - 1) / nw;
    int64_t item = rf_scalar64(gw * per);
    const int64_t item_end = rf_scalar64(item + per < nitems ? item + per : nitems);
    if (item >= item_end) return;
    int64_t u = 0, hi = nunits - 1;
    while (u < hi) {
        const int64_t mid = (u + hi + 1) >> 1;
        if (item_start[mid] <= item) u = mid;
        else hi = mid - 1;
    }
    u = rf_scalar64(u);
    for (; item < item_end; item++) {
        while (item_start[u + 1] <= item) u++;
        const RfScanArgs& L = work[u];
        const int64_t tile = item - item_start[u];
        const int64_t src = tile * TILE;
        const int64_t dst = rf_g(L.wave_dst)[tile], cnt = rf_g(L.wave_cnt)[tile];
        for (int64_t i = lane; i < cnt; i += 64) Q::move(L, src + i, dst + i);
    }
}
