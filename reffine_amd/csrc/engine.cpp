// HIP runtime layer (see engine.hpp).
#include "engine.hpp"

#include <atomic>
#include <map>
#include <unordered_map>

#include <hip/hiprtc.h>

#include <algorithm>
#include <list>
#include <cmath>
#include <cstring>
#include <functional>
#include <set>
#include <sstream>

namespace rf {

#include "rf_embed.inc"  // kRfAbiSrc, kRfDeviceSrc: kernel header texts

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

static Device g_dev;

Device& device() {
    if (g_dev.id < 0) throw HipError("rf_init has not been called");
    return g_dev;
}

Device& init_device(int id) {
    std::lock_guard<std::mutex> lk(g_dev.mu);
    if (g_dev.id == id) return g_dev;
    if (g_dev.id >= 0) throw HipError("library already bound to another device");
    int n = 0;
    RF_HIP(hipGetDeviceCount(&n));
    if (id < 0 || id >= n) throw HipError("no HIP device " + std::to_string(id));
    RF_HIP(hipSetDevice(id));
    hipDeviceProp_t prop;
    RF_HIP(hipGetDeviceProperties(&prop, id));
    std::string arch = prop.gcnArchName;
    if (arch.rfind("gfx950", 0) != 0)
        throw HipError("device " + std::to_string(id) + " is " + arch +
                       "; this library targets gfx950 (MI355X) only");
    g_dev.num_cus = prop.multiProcessorCount;
    g_dev.arch = "gfx950";
    RF_HIP(hipStreamCreateWithFlags(&g_dev.stream, hipStreamNonBlocking));
    noalloc_setup(id);
    g_dev.id = id;
    return g_dev;
}

// Device memory. Freed blocks up to 2 GiB are kept (8 GiB in total at
// most) and handed out again for the same size class: hipFree synchronises
// the device and hipMalloc costs tens of microseconds, and a repeated query
// allocates and frees its result table on every run. A block is freed only
// once nothing on the device uses it (callers free results after their run
// has finished).
namespace {
std::mutex g_mem_mu;
std::unordered_map<void*, size_t> g_live;  // block -> size class
std::multimap<size_t, void*> g_cache;      // free blocks by size class
size_t g_cached = 0;
constexpr size_t kCacheBlockMax = size_t(2) << 30;
constexpr size_t kCacheMax = size_t(8) << 30;

size_t size_class(size_t b) {
    if (b <= 256) return 256;
    if (b <= (size_t(1) << 20)) {
        size_t c = 256;
        while (c < b) c <<= 1;
        return c;
    }
    const size_t mb = size_t(1) << 20;
    return (b + mb - 1) / mb * mb;
}

void release_cache_locked() {
    for (auto& kv : g_cache) (void)hipFree(kv.second);
    g_cache.clear();
    g_cached = 0;
}
}  // namespace

void* dev_alloc(size_t bytes) {
    const size_t sz = size_class(bytes);
    std::lock_guard<std::mutex> lk(g_mem_mu);
    auto it = g_cache.find(sz);
    if (it != g_cache.end()) {
        void* p = it->second;
        g_cache.erase(it);
        g_cached -= sz;
        g_live[p] = sz;
        return p;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, sz);
    if (e != hipSuccess && !g_cache.empty()) {
        (void)hipGetLastError();
        release_cache_locked();
        e = hipMalloc(&p, sz);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw NoMem("device allocation of " + std::to_string(bytes) + " bytes failed: " +
                    hipGetErrorString(e));
    }
    g_live[p] = sz;
    return p;
}

void dev_free(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_mem_mu);
    auto it = g_live.find(p);
    if (it == g_live.end()) {
        (void)hipFree(p);
        return;
    }
    const size_t sz = it->second;
    g_live.erase(it);
    if (sz <= kCacheBlockMax && g_cached + sz <= kCacheMax) {
        g_cache.emplace(sz, p);
        g_cached += sz;
    } else {
        (void)hipFree(p);
    }
}

void dev_fill_u64(void* p, uint64_t v, int64_t n) {
    RF_HIP(aot_fill_u64(p, v, n, device().stream));
}

// ------------------------------------------------------------------ tables

static bool format_type(const std::string& f, Type& t) {
    static const std::map<std::string, BT> m = {
        {"c", BT::INT8},   {"C", BT::UINT8},  {"s", BT::INT16}, {"S", BT::UINT16},
        {"i", BT::INT32},  {"I", BT::UINT32}, {"l", BT::INT64}, {"L", BT::UINT64},
        {"f", BT::FLOAT32}, {"g", BT::FLOAT64}};
    auto it = m.find(f);
    if (it == m.end()) return false;
    t = Type(it->second);
    return true;
}

bool format_type_of(const std::string& f, Type& t) { return format_type(f, t); }

std::string arrow_format_of(const Type& t) {
    switch (t.bt) {
        case BT::BOOL: case BT::INT8: return "c";
        case BT::UINT8: return "C";
        case BT::INT16: return "s";
        case BT::UINT16: return "S";
        case BT::INT32: return "i";
        case BT::UINT32: return "I";
        case BT::INT64: case BT::IDX: return "l";
        case BT::UINT64: return "L";
        case BT::FLOAT32: return "f";
        case BT::FLOAT64: return "g";
        default: throw ArrowError("no Arrow format for type " + t.str());
    }
}

}  // namespace rf

static std::atomic<uint64_t> g_table_uid{1};
rf_table::rf_table()
    : uid(g_table_uid.fetch_add(1)), owned(std::make_shared<rf::OwnedBuffers>()) {}

rf::Type rf_table::type() const {
    std::vector<rf::Type> d;
    std::vector<rf::Enc> e;
    for (auto& c : cols) {
        d.push_back(c.type);
        e.push_back(c.enc);
    }
    return rf::Type(rf::BT::VECTOR, d, (int)dim, e);
}

namespace rf {

// Bits [bit0, bit0 + n) of an Arrow validity bitmap (LSB first) as a bitmap
// starting at bit 0: the source itself when bit0 is a whole byte, else a
// shifted host copy kept in `keep`.
static const void* bitmap_window(const void* bm, int64_t bit0, int64_t n,
                                 std::list<std::vector<uint8_t>>& keep) {
    if (!bm || n <= 0) return bm;
    const uint8_t* in = (const uint8_t*)bm;
    if (bit0 % 8 == 0) return in + bit0 / 8;
    const int64_t k = bit0 >> 3, last = (bit0 + n - 1) >> 3, nbytes = (n + 7) / 8;
    const int sh = (int)(bit0 & 7);
    keep.emplace_back((size_t)nbytes, (uint8_t)0);
    uint8_t* out = keep.back().data();
    for (int64_t j = 0; j < nbytes; j++) {
        const uint32_t lo = (uint32_t)in[k + j] >> sh;
        const uint32_t hi = (k + j + 1 <= last) ? (uint32_t)in[k + j + 1] << (8 - sh) : 0u;
        out[j] = (uint8_t)(lo | hi);
    }
    return out;
}

// Copies a record batch (Arrow C data interface, struct array) to HBM. Array
// offsets are honoured (a sliced batch uploads only its rows): a column's
// rows are [parent offset + column offset, + batch length); a run-end column
// keeps the runs that cover its window, with run ends cut to the window and
// made relative to its start.
rf_table* table_upload(const ArrowSchema* schema, const ArrowArray* array, int64_t dim) {
    if (!schema || !array) throw ArrowError("null schema or array");
    if (!schema->format || std::string(schema->format) != "+s")
        throw ArrowError("record batch must be a struct array (format \"+s\")");
    if (schema->n_children != array->n_children)
        throw ArrowError("schema/array child count mismatch");
    if (array->offset < 0 || array->length < 0) throw ArrowError("negative offset or length");
    if (dim < 1 || dim > schema->n_children) throw ArrowError("bad dim");
    Device& d = device();
    auto t = std::make_unique<rf_table>();
    t->dim = dim;
    struct Piece {
        const void* src;
        size_t bytes;
        size_t off;
        void** dst;
    };
    std::vector<Piece> pieces;
    std::list<std::vector<uint8_t>> keep;  // host copies alive until the copies finish
    size_t total = 0;
    auto add = [&](const void* src, size_t bytes, void** dst) {
        pieces.push_back({src, bytes, total, dst});
        total += (bytes + 255) & ~size_t(255);
        if (bytes == 0) total += 256;
    };
    const int64_t rows = array->length;
    t->cols.resize(schema->n_children);
    for (int64_t i = 0; i < schema->n_children; i++) {
        const ArrowSchema* cs = schema->children[i];
        const ArrowArray* ca = array->children[i];
        Column& c = t->cols[i];
        c.format = cs->format ? cs->format : "";
        const std::string col = "column " + std::to_string(i);
        const int64_t eo = array->offset + ca->offset;  // the column's first row
        if (ca->offset < 0 || ca->length < array->offset + rows)
            throw ArrowError(col + " is shorter than the batch");
        const ArrowArray* va = ca;
        const ArrowSchema* vs = cs;
        int64_t v0 = eo, nv = rows;  // values window
        if (c.format == "+r") {
            if (ca->n_children != 2 || cs->n_children != 2)
                throw ArrowError("run-end column needs run_ends and values children");
            if (std::string(cs->children[0]->format) != "i")
                throw ArrowError("run ends must be int32");
            const ArrowArray* re = ca->children[0];
            va = ca->children[1];
            vs = cs->children[1];
            c.enc = Enc::RUNEND;
            c.logical_length = rows;
            const int32_t* ends = (const int32_t*)re->buffers[1] + re->offset;
            const int64_t nr = re->length;
            // Runs covering logical rows [eo, eo + rows): r0 is the first run
            // ending after eo, r1 the first ending at or after eo + rows.
            const int32_t* p0 = std::upper_bound(ends, ends + nr, (int32_t)eo);
            const int64_t r0 = p0 - ends;
            const int64_t r1 = std::lower_bound(ends, ends + nr, (int32_t)(eo + rows)) - ends;
            const int64_t nruns = rows == 0 ? 0 : std::min<int64_t>(r1, nr - 1) - r0 + 1;
            if (rows > 0 && (r0 >= nr || r1 >= nr))
                throw ArrowError(col + ": run ends do not cover the batch");
            if (eo == 0 && nruns == nr && (nr == 0 || ends[nr - 1] == rows)) {
                add(ends, (size_t)nr * 4, &c.run_ends);
            } else {
                keep.emplace_back((size_t)std::max<int64_t>(nruns, 0) * 4, (uint8_t)0);
                int32_t* cut = (int32_t*)keep.back().data();
                for (int64_t k = 0; k < nruns; k++)
                    cut[k] = (int32_t)(std::min<int64_t>(ends[r0 + k], eo + rows) - eo);
                add(cut, (size_t)nruns * 4, &c.run_ends);
            }
            v0 = va->offset + r0;
            nv = nruns;
            if (va->offset < 0 || va->length < r0 + nruns)
                throw ArrowError(col + ": fewer run values than runs");
        }
        if (!format_type(vs->format ? vs->format : "", c.type))
            throw ArrowError("unsupported Arrow format '" + std::string(vs->format) + "'");
        c.length = nv;
        if (c.enc == Enc::FLAT) c.logical_length = c.length;
        if (va->n_buffers < 2) throw ArrowError("primitive array without a values buffer");
        const size_t w = (size_t)c.type.width();
        add((const char*)va->buffers[1] + (size_t)v0 * w, (size_t)nv * w, &c.values);
        // A window of a column with nulls may hold none; the bitmap goes up
        // either way (null_count -1 means not computed).
        c.null_count = (v0 == 0 && nv == va->length) ? va->null_count : -1;
        if (va->buffers[0] && va->null_count != 0) {
            const void* bm = bitmap_window(va->buffers[0], v0, nv, keep);
            add(bm, ((size_t)nv + 7) / 8, &c.validity);
        }
    }
    char* arena = (char*)dev_alloc(total);
    t->owned->p.push_back(arena);
    for (auto& p : pieces) {
        *p.dst = arena + p.off;
        if (p.bytes && p.src)
            RF_HIP(hipMemcpyAsync(arena + p.off, p.src, p.bytes, hipMemcpyHostToDevice, d.stream));
        // Bitmaps are read as uint16 words: zero the padding byte(s).
        size_t padded = (p.bytes + 255) & ~size_t(255);
        if (padded > p.bytes)
            RF_HIP(hipMemsetAsync(arena + p.off + p.bytes, 0, padded - p.bytes, d.stream));
    }
    RF_HIP(hipStreamSynchronize(d.stream));
    return t.release();
}

rf_table* table_wrap(int64_t dim, int64_t n_cols, int64_t n_rows, const char* const* formats,
                     void* const* values, void* const* validity) {
    if (n_cols < 1 || dim < 1 || dim > n_cols) throw ArrowError("bad table shape");
    auto t = std::make_unique<rf_table>();
    t->dim = dim;
    t->cols.resize(n_cols);
    for (int64_t i = 0; i < n_cols; i++) {
        Column& c = t->cols[i];
        c.format = formats[i];
        if (!format_type(c.format, c.type))
            throw ArrowError("unsupported Arrow format '" + c.format + "'");
        c.length = c.logical_length = n_rows;
        c.values = values[i];
        c.validity = validity ? validity[i] : nullptr;
        if ((uintptr_t)c.values % 16 != 0) throw ArrowError("device column not 16-byte aligned");
    }
    return t.release();
}

static uint64_t next_pow2(uint64_t x) {
    uint64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

// Column 0's key range (see rf_table::key_range_state); false when column
// 0 is not a flat int64 column without nulls, or has no rows.
static bool table_key_range(rf_table* t, int64_t& kmin, int64_t& kmax) {
    if (t->key_range_state == 0) {
        t->key_range_state = -1;
        const Column* c = t->cols.empty() ? nullptr : &t->cols[0];
        if (c && c->type == types::I64 && c->enc == Enc::FLAT && !c->validity && c->length > 0) {
            Device& d = device();
            uint64_t* small = (uint64_t*)dev_alloc(32);
            const uint64_t init[4] = {~0ull, 0ull, (uint64_t)c->length, 0ull};
            RF_HIP(hipMemcpyAsync(small, init, 32, hipMemcpyHostToDevice, d.stream));
            RF_HIP(aot_minmax_i64((const int64_t*)c->values, (const int64_t*)(small + 2), c->length,
                                  small, d.stream));
            uint64_t mm[2];
            RF_HIP(hipMemcpyAsync(mm, small, 16, hipMemcpyDeviceToHost, d.stream));
            RF_HIP(hipStreamSynchronize(d.stream));
            dev_free(small);
            t->key_min = (int64_t)(mm[0] ^ 0x8000000000000000ull);
            t->key_max = (int64_t)(mm[1] ^ 0x8000000000000000ull);
            t->key_range_state = 1;
        }
    }
    if (t->key_range_state != 1) return false;
    kmin = t->key_min;
    kmax = t->key_max;
    return true;
}

void table_build_index(rf_table* t) {
    if (t->cols.empty()) throw ArrowError("empty table");
    // Over the values of column 0: its rows, or for a run-end column its run
    // values (key -> run), as ArrowTable2::build_index indexes
    // (include/reffine/arrow/table.h:107-138).
    const Column& c = t->cols[0];
    if (c.type != types::I64)
        throw ArrowError("index needs an int64 column 0 (include/reffine/arrow/table.h:107-138)");
    Device& d = device();
    if (t->has_index) return;
    // Keys dense in their range (all valid, at most 32 key slots per row
    // plus 2^20): a rank map -- 1.5 bits per slot, small enough to stay in
    // the caches, where a probe of the hash slots is a random HBM read.
    if (!c.validity && c.length > 0 && c.length < ((int64_t)1 << 31)) {
        uint64_t* small = (uint64_t*)dev_alloc(32);
        const uint64_t init[4] = {~0ull, 0ull, (uint64_t)c.length, 0ull};
        RF_HIP(hipMemcpyAsync(small, init, 32, hipMemcpyHostToDevice, d.stream));
        RF_HIP(aot_minmax_i64((const int64_t*)c.values, (const int64_t*)(small + 2), c.length,
                              small, d.stream));
        RF_HIP(aot_not_ascending((const int64_t*)c.values, c.length, small + 3, d.stream));
        uint64_t mm[4];
        RF_HIP(hipMemcpyAsync(mm, small, 32, hipMemcpyDeviceToHost, d.stream));
        RF_HIP(hipStreamSynchronize(d.stream));
        dev_free(small);
        const int64_t kmin = (int64_t)(mm[0] ^ 0x8000000000000000ull);
        const int64_t kmax = (int64_t)(mm[1] ^ 0x8000000000000000ull);
        const uint64_t range = (uint64_t)kmax - (uint64_t)kmin + 1;
        if (range != 0 && range <= 32 * (uint64_t)c.length + ((uint64_t)1 << 20)) {
            const int64_t words = (int64_t)((range + 63) / 64);
            void* bits = dev_alloc((size_t)words * 8);
            t->owned->p.push_back(bits);
            void* cum = dev_alloc((size_t)aot_rank_cum_entries(words) * 4);
            t->owned->p.push_back(cum);
            dev_fill_u64(bits, 0, words);
            RF_HIP(aot_rank_build((const int64_t*)c.values, c.length, kmin, bits, words, cum,
                                  d.stream));
            void* rows = nullptr;
            if (mm[3] != 0) {  // not strictly ascending: the first row of each key
                rows = dev_alloc((size_t)c.length * 8);
                t->owned->p.push_back(rows);
                dev_fill_u64(rows, (uint64_t)INT64_MAX, c.length);
                RF_HIP(aot_rank_rows((const int64_t*)c.values, c.length, kmin, bits, cum,
                                     (int64_t*)rows, d.stream));
            }
            void* pack = dev_alloc((size_t)words * 16);
            t->owned->p.push_back(pack);
            RF_HIP(aot_rank_pack(bits, cum, words, pack, d.stream));
            RF_HIP(hipStreamSynchronize(d.stream));
            t->index.dense_pack = pack;
            t->index.dense_bits = bits;
            t->index.dense_cum = cum;
            t->index.dense_rows = rows;
            t->index.dense_base = kmin;
            t->index.dense_len = range;
            t->has_index = true;
            return;
        }
    }
    const uint64_t cap = next_pow2(std::max<uint64_t>(16, 2 * (uint64_t)c.length));
    void* keys = dev_alloc(cap * 8);
    t->owned->p.push_back(keys);
    void* rows = dev_alloc(cap * 8);
    t->owned->p.push_back(rows);
    void* mk = dev_alloc(8);
    dev_fill_u64(keys, (uint64_t)RF_EMPTY_KEY, (int64_t)cap);
    dev_fill_u64(rows, (uint64_t)INT64_MAX, (int64_t)cap);
    dev_fill_u64(mk, (uint64_t)INT64_MAX, 1);
    RF_HIP(aot_index_build((const int64_t*)c.values, (const uint16_t*)c.validity, c.length, keys,
                           rows, cap - 1, mk, d.stream));
    int64_t mkr = 0;
    RF_HIP(hipMemcpyAsync(&mkr, mk, 8, hipMemcpyDeviceToHost, d.stream));
    RF_HIP(hipStreamSynchronize(d.stream));
    dev_free(mk);
    t->index.keys = keys;
    t->index.rows = rows;
    t->index.mask = cap - 1;
    t->index.min_key_row = (mkr == INT64_MAX) ? -1 : mkr;
    t->has_index = true;
}

RfProbeDev probe_of(const rf_table* t) {
    RfProbeDev h{};
    h.keys = (const int64_t*)t->index.keys;
    h.rows = (const int64_t*)t->index.rows;
    h.mask = t->index.mask;
    h.min_key_row = t->index.min_key_row;
    h.dense_bits = (const uint64_t*)t->index.dense_bits;
    h.dense_cum = (const uint32_t*)t->index.dense_cum;
    h.dense_base = t->index.dense_base;
    h.dense_len = t->index.dense_len;
    h.dense_rows = (const int64_t*)t->index.dense_rows;
    h.dense_pack = (const uint64_t*)t->index.dense_pack;
    return h;
}

// -------------------------------------------------------------------- JIT

static std::mutex g_jit_mu;
static std::map<std::string, std::unique_ptr<Module>> g_modules;

static std::string full_source(const KernelSource& ks) {
    return std::string("#include \"rf_device.h\"\n") + ks.code;
}

std::vector<char> jit_compile(const KernelSource& ks) {
    const std::string src = full_source(ks);
    const char* headers[] = {kRfAbiSrc, kRfDeviceSrc};
    const char* names[] = {"rf_abi.h", "rf_device.h"};
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "rf_query.hip", 2, headers, names) !=
        HIPRTC_SUCCESS)
        throw HipError("hiprtcCreateProgram failed");
    // gfx950 only; no FMA contraction, no packed FP32.
    const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                          "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang",
                          "-packed-fp32-ops"};
    hiprtcResult r = hiprtcCompileProgram(prog, sizeof(opts) / sizeof(opts[0]), opts);
    size_t log_size = 0;
    hiprtcGetProgramLogSize(prog, &log_size);
    std::string log(log_size, '\0');
    if (log_size) hiprtcGetProgramLog(prog, &log[0]);
    if (r != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        throw HipError("hiprtc compile failed: " + log + "\n--- source ---\n" + ks.code);
    }
    size_t code_size = 0;
    hiprtcGetCodeSize(prog, &code_size);
    std::vector<char> code(code_size);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    return code;
}

Module& jit_module(const KernelSource& ks) {
    device();
    const std::string src = full_source(ks);
    std::lock_guard<std::mutex> lk(g_jit_mu);
    auto it = g_modules.find(src);
    if (it != g_modules.end()) return *it->second;
    std::vector<char> code = jit_compile(ks);
    auto m = std::make_unique<Module>();
    RF_HIP(hipModuleLoadData(&m->mod, code.data()));
    RF_HIP(hipModuleGetFunction(&m->fn, m->mod, ks.entry.c_str()));
    if (ks.entry == "rf_seg_fold" && ks.alt_cap > 0)
        RF_HIP(hipModuleGetFunction(&m->seg_long, m->mod, "rf_seg_fold_long"));
    if (ks.entry == "rf_seg_fold")
        RF_HIP(hipModuleGetFunction(&m->seg_wave, m->mod, "rf_seg_fold_wave"));
    if (ks.entry == "rf_gj_fold") {
        RF_HIP(hipModuleGetFunction(&m->fold_b, m->mod, "rf_gj_fold_b"));
        RF_HIP(hipModuleGetFunction(&m->fold_w, m->mod, "rf_gj_fold_w"));
        RF_HIP(hipModuleGetFunction(&m->fold_wb, m->mod, "rf_gj_fold_wb"));
        if (ks.code.find("void rf_gj_agg(") != std::string::npos)
            RF_HIP(hipModuleGetFunction(&m->agg, m->mod, "rf_gj_agg"));
        if (ks.code.find("void rf_gj_lds_done(") != std::string::npos)
            RF_HIP(hipModuleGetFunction(&m->lds_done, m->mod, "rf_gj_lds_done"));
    }
    if (ks.entry == "rf_qu_count") RF_HIP(hipModuleGetFunction(&m->emit, m->mod, "rf_qu_emit"));
    if (ks.entry == "rf_q") {
        RF_HIP(hipModuleGetFunction(&m->gather, m->mod, "rf_q_gather"));
        RF_HIP(hipModuleGetFunction(&m->bscan, m->mod, "rf_qb"));
        RF_HIP(hipModuleGetFunction(&m->bdone, m->mod, "rf_qb_done"));
        RF_HIP(hipModuleGetFunction(&m->bgather, m->mod, "rf_qb_gather"));
        int nb = 0;
        RF_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nb, m->bscan, ks.tune.block, 0));
        m->bscan_blocks_per_cu = std::max(1, nb);
    }
    int nb = 0;
    RF_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nb, m->fn, ks.tune.block, 0));
    m->max_blocks_per_cu = std::max(1, nb);
    Module& ref = *m;
    g_modules[src] = std::move(m);
    return ref;
}

// ------------------------------------------------------- host evaluation

namespace {

struct EvalCtx {
    const Plan* plan;
    const rf_arg* args;
    int64_t n_args;
    std::map<std::string, HostVal> comps;  // component symbol -> value
    std::map<std::string, N> tbl;
};

rf_table* arg_table(const EvalCtx& c, int i) {
    if (i < 0 || i >= c.n_args || c.args[i].kind != 0 || !c.args[i].u.table)
        throw IRError("argument " + std::to_string(i) + " must be a table");
    return c.args[i].u.table;
}

HostVal hv_int(const Type& t, int64_t v) {
    HostVal h;
    h.type = t;
    h.i = v;
    return h;
}
HostVal hv_float(const Type& t, double v) {
    HostVal h;
    h.type = t;
    h.is_float = true;
    h.f = v;
    return h;
}

int64_t wrap_int(const Type& t, int64_t v) {
    switch (t.bt) {
        case BT::BOOL: return v != 0;
        case BT::INT8: return (int8_t)v;
        case BT::UINT8: return (uint8_t)v;
        case BT::INT16: return (int16_t)v;
        case BT::UINT16: return (uint16_t)v;
        case BT::INT32: return (int32_t)v;
        case BT::UINT32: return (uint32_t)v;
        default: return v;
    }
}

HostVal convert(const HostVal& v, const Type& t) {
    if (t.is_float()) {
        double x = v.is_float ? v.f
                              : (v.type.bt == BT::UINT64 ? (double)(uint64_t)v.i : (double)v.i);
        if (t.bt == BT::FLOAT32) x = (float)x;
        return hv_float(t, x);
    }
    int64_t x;
    if (v.is_float) {
        x = t.bt == BT::UINT64 ? (int64_t)(uint64_t)v.f : (int64_t)v.f;
        if (t.bt == BT::BOOL) x = v.f != 0;
    } else {
        x = v.i;
    }
    return hv_int(t, wrap_int(t, x));
}

HostVal eval(const N& e, EvalCtx& c);

HostVal eval_leaf(const N& e, EvalCtx& c) {
    const std::string n = e->name;
    if (n.rfind("$arg.", 0) == 0) {
        int i = std::stoi(n.substr(5));
        if (i >= c.n_args) throw IRError("missing argument " + std::to_string(i));
        const rf_arg& a = c.args[i];
        if (a.kind == 1) return convert(hv_int(types::I64, a.u.i64), e->type);
        if (a.kind == 2) return convert(hv_float(types::F64, a.u.f64), e->type);
        throw IRError("argument " + std::to_string(i) + " must be a scalar");
    }
    if (n.rfind("$len.", 0) == 0) {
        auto dot = n.find('.', 5);
        int t = std::stoi(n.substr(5, dot - 5));
        int col = std::stoi(n.substr(dot + 1));
        rf_table* tb = arg_table(c, t);
        if (col >= (int)tb->cols.size()) throw IRError("length of a missing column");
        return hv_int(types::IDX_T, tb->cols[col].length);
    }
    throw IRError("leaf " + n + " is not host-evaluable");
}

HostVal eval(const N& e, EvalCtx& c) {
    switch (e->kind) {
        case K::Const:
            if (e->type.is_float()) return convert(hv_float(types::F64, e->val), e->type);
            if (e->type.is_unsigned()) return hv_int(e->type, wrap_int(e->type, (int64_t)(uint64_t)e->val));
            return hv_int(e->type, wrap_int(e->type, (int64_t)e->val));
        case K::Sym: {
            if (!leaf_name(e).empty()) return eval_leaf(e, c);
            auto cm = c.comps.find(e->name);
            if (cm != c.comps.end()) return cm->second;
            for (size_t i = 0; i < c.plan->inputs.size(); i++) {
                if (c.plan->inputs[i]->name == e->name) {
                    if (e->type.is_vector()) throw IRError("vector input in a host expression");
                    N l = mk_sym("$arg." + std::to_string(i), e->type);
                    return eval_leaf(l, c);
                }
            }
            auto t = c.tbl.find(e->name);
            if (t != c.tbl.end()) return eval(t->second, c);
            throw IRError("undefined symbol " + e->name);
        }
        case K::Length: {
            N v = e->args[0];
            while (v->kind == K::Sym) {
                bool found = false;
                for (size_t i = 0; i < c.plan->inputs.size(); i++)
                    if (c.plan->inputs[i]->name == v->name) {
                        rf_table* tb = arg_table(c, (int)i);
                        if (e->idx >= (int64_t)tb->cols.size()) throw IRError("length column");
                        return hv_int(types::IDX_T, tb->cols[e->idx].length);
                    }
                auto t = c.tbl.find(v->name);
                if (t == c.tbl.end()) break;
                v = t->second;
                found = true;
                (void)found;
            }
            throw Unsupported("length of a computed vector on the host");
        }
        case K::Cast:
            return convert(eval(e->args[0], c), e->type);
        case K::Get: {
            HostVal v = eval(e->args[0], c);
            if (!v.type.is_struct()) return v;
            return v.fields.at(e->idx);
        }
        case K::New: {
            HostVal h;
            h.type = e->type;
            for (auto& a : e->args) h.fields.push_back(eval(a, c));
            return h;
        }
        case K::Select: {
            HostVal cv = eval(e->args[0], c);
            return (cv.is_float ? cv.f != 0 : cv.i != 0) ? eval(e->args[1], c)
                                                         : eval(e->args[2], c);
        }
        case K::Nary: {
            std::vector<HostVal> a;
            for (auto& x : e->args) a.push_back(eval(x, c));
            const Type& at = e->args[0]->type;
            bool fl = at.is_float();
            auto F = [&](int i) { return convert(a[i], types::F64).f; };
            auto I = [&](int i) { return a[i].is_float ? (int64_t)a[i].f : a[i].i; };
            auto B = [&](bool b) { return hv_int(types::BOOL_T, b ? 1 : 0); };
            switch (e->op) {
                case MathOp::ADD: return fl ? convert(hv_float(types::F64, F(0) + F(1)), e->type) : hv_int(e->type, wrap_int(e->type, (int64_t)((uint64_t)I(0) + (uint64_t)I(1))));
                case MathOp::SUB: return fl ? convert(hv_float(types::F64, F(0) - F(1)), e->type) : hv_int(e->type, wrap_int(e->type, (int64_t)((uint64_t)I(0) - (uint64_t)I(1))));
                case MathOp::MUL: return fl ? convert(hv_float(types::F64, F(0) * F(1)), e->type) : hv_int(e->type, wrap_int(e->type, (int64_t)((uint64_t)I(0) * (uint64_t)I(1))));
                case MathOp::DIV:
                    if (fl) return convert(hv_float(types::F64, F(0) / F(1)), e->type);
                    if (I(1) == 0) throw IRError("integer division by zero");
                    return hv_int(e->type, wrap_int(e->type, I(0) / I(1)));
                case MathOp::MOD:
                    if (I(1) == 0) throw IRError("integer modulo by zero");
                    return hv_int(e->type, wrap_int(e->type, I(0) % I(1)));
                case MathOp::MAX: return fl ? ((F(0) < F(1)) ? a[1] : a[0]) : ((I(0) < I(1)) ? a[1] : a[0]);
                case MathOp::MIN: return fl ? ((F(1) < F(0)) ? a[1] : a[0]) : ((I(1) < I(0)) ? a[1] : a[0]);
                case MathOp::ABS: return fl ? convert(hv_float(types::F64, std::fabs(F(0))), e->type) : hv_int(e->type, I(0) < 0 ? -I(0) : I(0));
                case MathOp::NEG: return fl ? convert(hv_float(types::F64, -F(0)), e->type) : hv_int(e->type, wrap_int(e->type, -I(0)));
                case MathOp::SQRT: return convert(hv_float(types::F64, std::sqrt(F(0))), e->type);
                case MathOp::POW: return convert(hv_float(types::F64, std::pow(F(0), F(1))), e->type);
                case MathOp::CEIL: return convert(hv_float(types::F64, std::ceil(F(0))), e->type);
                case MathOp::FLOOR: return convert(hv_float(types::F64, std::floor(F(0))), e->type);
                case MathOp::LT: return B(fl ? F(0) < F(1) : I(0) < I(1));
                case MathOp::LTE: return B(fl ? F(0) <= F(1) : I(0) <= I(1));
                case MathOp::GT: return B(fl ? F(0) > F(1) : I(0) > I(1));
                case MathOp::GTE: return B(fl ? F(0) >= F(1) : I(0) >= I(1));
                case MathOp::EQ: return B(fl ? F(0) == F(1) : I(0) == I(1));
                case MathOp::NOT: return B(I(0) == 0);
                case MathOp::AND: return B(I(0) != 0 && I(1) != 0);
                case MathOp::OR: return B(I(0) != 0 || I(1) != 0);
            }
            break;
        }
        default:
            break;
    }
    throw Unsupported("expression not host-evaluable: " + to_text(e));
}

int64_t eval_idx(const N& e, EvalCtx& c) {
    HostVal v = eval(e, c);
    return v.is_float ? (int64_t)v.f : v.i;
}

HostVal combine_host(AggOp op, const HostVal& a, const HostVal& b, const Type& t) {
    bool fl = t.is_float();
    HostVal r;
    switch (op) {
        case AggOp::ADD:
            r = fl ? hv_float(t, a.f + b.f) : hv_int(t, (int64_t)((uint64_t)a.i + (uint64_t)b.i));
            break;
        case AggOp::MUL:
            r = fl ? hv_float(t, a.f * b.f) : hv_int(t, (int64_t)((uint64_t)a.i * (uint64_t)b.i));
            break;
        case AggOp::MIN:
            r = fl ? ((b.f < a.f) ? b : a) : ((b.i < a.i) ? b : a);
            break;
        case AggOp::MAX:
            r = fl ? ((a.f < b.f) ? b : a) : ((a.i < b.i) ? b : a);
            break;
        case AggOp::AND: r = hv_int(t, (a.i != 0 && b.i != 0) ? 1 : 0); break;
        case AggOp::OR: r = hv_int(t, (a.i != 0 || b.i != 0) ? 1 : 0); break;
    }
    if (fl && t.bt == BT::FLOAT32) r.f = (float)r.f;
    if (!fl) r.i = wrap_int(t, r.i);
    r.type = t;
    return r;
}

HostVal read_prim(const Type& t, const char* p) {
    switch (t.bt) {
        case BT::BOOL: return hv_int(t, *(const uint8_t*)p != 0);
        case BT::INT8: return hv_int(t, *(const int8_t*)p);
        case BT::UINT8: return hv_int(t, *(const uint8_t*)p);
        case BT::INT16: return hv_int(t, *(const int16_t*)p);
        case BT::UINT16: return hv_int(t, *(const uint16_t*)p);
        case BT::INT32: return hv_int(t, *(const int32_t*)p);
        case BT::UINT32: return hv_int(t, *(const uint32_t*)p);
        case BT::INT64: case BT::IDX: return hv_int(t, *(const int64_t*)p);
        case BT::UINT64: return hv_int(t, (int64_t) * (const uint64_t*)p);
        case BT::FLOAT32: return hv_float(t, *(const float*)p);
        case BT::FLOAT64: return hv_float(t, *(const double*)p);
        default: throw IRError("bad primitive");
    }
}

void write_prim(const HostVal& v, const Type& t, char* p) {
    HostVal c = convert(v, t);
    switch (t.bt) {
        case BT::BOOL: *(uint8_t*)p = c.i != 0; break;
        case BT::INT8: *(int8_t*)p = (int8_t)c.i; break;
        case BT::UINT8: *(uint8_t*)p = (uint8_t)c.i; break;
        case BT::INT16: *(int16_t*)p = (int16_t)c.i; break;
        case BT::UINT16: *(uint16_t*)p = (uint16_t)c.i; break;
        case BT::INT32: *(int32_t*)p = (int32_t)c.i; break;
        case BT::UINT32: *(uint32_t*)p = (uint32_t)c.i; break;
        case BT::INT64: case BT::IDX: *(int64_t*)p = c.i; break;
        case BT::UINT64: *(uint64_t*)p = (uint64_t)c.i; break;
        case BT::FLOAT32: *(float*)p = (float)c.f; break;
        case BT::FLOAT64: *(double*)p = c.f; break;
        default: throw IRError("bad primitive");
    }
}

HostVal read_value(const Type& t, const char* src) {
    if (t.is_struct()) {
        HostVal v;
        v.type = t;
        size_t off = 0;
        for (size_t i = 0; i < t.dtypes.size(); i++) {
            size_t sz, al;
            c_layout(t.dtypes[i], sz, al);
            off = (off + al - 1) / al * al;
            v.fields.push_back(read_value(t.dtypes[i], src + off));
            off += sz;
        }
        return v;
    }
    if (t.is_vector()) {
        HostVal v;
        v.type = t;
        v.table = *(void* const*)src;
        return v;
    }
    return read_prim(t, src);
}

void write_value(const HostVal& v, const Type& t, char* dst) {
    if (t.is_struct()) {
        size_t off = 0;
        for (size_t i = 0; i < t.dtypes.size(); i++) {
            size_t sz, al;
            c_layout(t.dtypes[i], sz, al);
            off = (off + al - 1) / al * al;
            write_value(v.fields.at(i), t.dtypes[i], dst + off);
            off += sz;
        }
        return;
    }
    if (t.is_vector()) {
        *(void**)dst = v.table;
        return;
    }
    write_prim(v, t, dst);
}

}  // namespace

void c_layout(const Type& t, size_t& size, size_t& align) {
    if (t.is_struct()) {
        size_t off = 0, al = 1;
        for (auto& d : t.dtypes) {
            size_t s, a;
            c_layout(d, s, a);
            off = (off + a - 1) / a * a + s;
            al = std::max(al, a);
        }
        size = (off + al - 1) / al * al;
        align = al;
        return;
    }
    if (t.is_vector() || t.bt == BT::PTR) {
        size = align = 8;
        return;
    }
    size = align = (size_t)t.width();
}

// ----------------------------------------------------------- execution

static std::string null_signature(const std::vector<N>& inputs, const rf_arg* args,
                                  int64_t n_args) {
    std::string s;
    for (size_t i = 0; i < inputs.size(); i++) {
        if (!inputs[i]->type.is_vector()) continue;
        if ((int64_t)i >= n_args || args[i].kind != 0 || !args[i].u.table)
            throw IRError("argument " + std::to_string(i) + " must be a table");
        const rf_table* t = args[i].u.table;
        // idx columns are stored as int64 (a computed table's iterators).
        std::function<Type(Type)> norm = [&](Type x) {
            if (x.bt == BT::IDX) x.bt = BT::INT64;
            for (auto& d : x.dtypes) d = norm(d);
            return x;
        };
        const Type& want = inputs[i]->type;
        Type have = t->type();
        if (norm(have) != norm(want))
            throw IRError("argument " + std::to_string(i) + " has type " + have.str() +
                          ", the function expects " + want.str());
        for (auto& c : t->cols) s += c.validity ? '1' : '0';
        s += '|';
    }
    return s;
}

static std::vector<std::vector<bool>> nullable_of(size_t n_inputs, const rf_arg* args,
                                                  const std::vector<N>& inputs) {
    std::vector<std::vector<bool>> nl(n_inputs);
    for (size_t i = 0; i < n_inputs; i++) {
        if (!inputs[i]->type.is_vector()) continue;
        for (auto& c : args[i].u.table->cols) nl[i].push_back(c.validity != nullptr);
    }
    return nl;
}

// The run stage's inputs: the function's, with T's own run-end type.
static std::vector<N> run_stage_inputs(const Plan& plan) {
    std::vector<N> in = plan.inputs;
    in[plan.runs->table] = plan.runs->input;
    return in;
}

// The run-reduction kernel of a run stage, by the inputs' nullability.
static Spec& get_run_spec(rf_query* q, const rf_arg* args, int64_t n_args) {
    const std::vector<N> inputs = run_stage_inputs(q->plan);
    std::string sig = null_signature(inputs, args, n_args);
    // And by each probed table's index form: a packed rank map over
    // ascending keys ('p') is probed without branches.
    std::vector<bool> pack;
    for (const Probe& pr : q->plan.runs->seg.sub.probes) {
        rf_table* pt = args[pr.table].u.table;
        if (!pt->has_index) table_build_index(pt);
        const bool p = pt->index.dense_pack && !pt->index.dense_rows && pt->index.dense_len > 0;
        pack.push_back(p);
        sig += p ? 'p' : 'h';
    }
    auto it = q->run_specs.find(sig);
    if (it != q->run_specs.end()) return it->second;
    Spec sp;
    sp.ks = gen_seg_fold(q->plan.runs->seg, q->plan, nullable_of(inputs.size(), args, inputs),
                         pack);
    sp.mod = &jit_module(sp.ks);
    return q->run_specs.emplace(sig, std::move(sp)).first->second;
}

// The scan kernel's key: the inputs' nullability, then each probed table's
// index form (a packed rank map over ascending keys, 'p', is probed without
// branches, so a tile's rows issue their probes together). `build` builds a
// missing index first (a table without one keys as 'h').
static std::string scan_signature(rf_query* q, const rf_arg* args, int64_t n_args, bool build,
                                  std::vector<bool>* pack) {
    std::string sig = null_signature(q->plan.inputs, args, n_args);
    if (q->plan.seg || q->plan.source != Plan::TABLE_ROWS) return sig;
    for (const Probe& pr : q->plan.probes) {
        bool p = false;
        if (!pr.external && pr.table >= 0 && pr.table < n_args) {
            rf_table* pt = args[pr.table].u.table;
            if (build && !pt->has_index) table_build_index(pt);
            p = pt->has_index && pt->index.dense_pack && !pt->index.dense_rows &&
                pt->index.dense_len > 0;
        }
        if (pack) pack->push_back(p);
        sig += p ? 'p' : 'h';
    }
    return sig;
}

Spec* find_spec(rf_query* q, const rf_arg* args, int64_t n_args) {
    auto it = q->specs.find(scan_signature(q, args, n_args, false, nullptr));
    return it == q->specs.end() ? nullptr : &it->second;
}

static Spec& get_spec(rf_query* q, const rf_arg* args, int64_t n_args) {
    std::vector<bool> pack;
    std::string sig = scan_signature(q, args, n_args, true, &pack);
    auto it = q->specs.find(sig);
    if (it != q->specs.end()) return it->second;
    std::vector<std::vector<bool>> nl(q->plan.inputs.size());
    for (size_t i = 0; i < q->plan.inputs.size(); i++) {
        if (!q->plan.inputs[i]->type.is_vector()) continue;
        for (auto& c : args[i].u.table->cols) nl[i].push_back(c.validity != nullptr);
    }
    Spec sp;
    sp.ks = q->plan.seg ? gen_seg_fold(*q->plan.seg, q->plan, nl)
                        : gen_scan(q->plan, nl, q->tune, q->det, pack);
    sp.mod = &jit_module(sp.ks);
    auto r = q->specs.emplace(sig, std::move(sp));
    return r.first->second;
}

}  // namespace rf

rf_query::~rf_query() {
    rf::dev_free(partials);
    rf::dev_free(result);
    rf::dev_free(counters);
    rf::dev_free(wave_arrays);
    rf::dev_free(out_count);
    for (void* p : outs) rf::dev_free(p);
    for (void* p : stages) rf::dev_free(p);
    if (host_stage) (void)hipHostFree(host_stage);
}

namespace rf {

OwnedBuffers::~OwnedBuffers() {
    for (void* x : p) dev_free(x);
}

void table_free(rf_table* t) {
    delete t;  // its buffers go with the last holder
}

static void fill_params(const KernelSource& ks, EvalCtx& c, RfParam* out) {
    for (size_t i = 0; i < ks.params.size(); i++) {
        const ParamSlot& p = ks.params[i];
        HostVal v;
        if (p.kind == ParamSlot::CONST) {
            // A constant's value as host evaluation gives it (K::Const).
            if (p.type.is_float()) out[i].f = p.value;
            else if (p.type.is_unsigned()) out[i].i = wrap_int(p.type, (int64_t)(uint64_t)p.value);
            else out[i].i = wrap_int(p.type, (int64_t)p.value);
            continue;
        } else if (p.kind == ParamSlot::ARG) {
            v = eval_leaf(mk_sym("$arg." + std::to_string(p.arg), p.type), c);
        } else {
            v = eval_leaf(mk_sym("$len." + std::to_string(p.arg) + "." + std::to_string(p.col),
                                 types::IDX_T), c);
        }
        if (p.type.is_float()) out[i].f = convert(v, types::F64).f;
        else out[i].i = convert(v, types::I64).i;
    }
}

void check_args(const Plan& plan, const rf_arg* args, int64_t n_args) {
    if (n_args != (int64_t)plan.inputs.size())
        throw IRError("function '" + plan.func_name + "' takes " +
                      std::to_string(plan.inputs.size()) + " arguments, got " +
                      std::to_string(n_args));
    for (size_t i = 0; i < plan.inputs.size(); i++) {
        bool vec = plan.inputs[i]->type.is_vector();
        if (vec != (args[i].kind == 0))
            throw IRError("argument " + std::to_string(i) + (vec ? " must be a table" : " must be a scalar"));
    }
}

// Output column descriptor (Arrow format, storage type) of `count` rows.
static Column out_column(const Type& ty, int64_t count) {
    Column col;
    col.type = ty;
    switch (ty.bt) {
        case BT::BOOL: case BT::INT8: col.format = "c"; break;
        case BT::UINT8: col.format = "C"; break;
        case BT::INT16: col.format = "s"; break;
        case BT::UINT16: col.format = "S"; break;
        case BT::INT32: col.format = "i"; break;
        case BT::UINT32: col.format = "I"; break;
        case BT::INT64: case BT::IDX: col.format = "l"; col.type = types::I64; break;
        case BT::UINT64: col.format = "L"; break;
        case BT::FLOAT32: col.format = "f"; break;
        case BT::FLOAT64: col.format = "g"; break;
        default: throw IRError("output column type");
    }
    if (ty.bt == BT::BOOL) col.type = types::I8;
    col.length = col.logical_length = count;
    return col;
}

// A device table of `count` rows with fresh columns of `types` (one arena).
rf_table* alloc_table(const std::vector<Type>& types, int64_t count, int64_t dim) {
    auto t = std::make_unique<rf_table>();
    t->dim = dim;
    size_t total = 0;
    std::vector<size_t> offs;
    for (auto& ty : types) {
        offs.push_back(total);
        total += ((size_t)std::max<int64_t>(count, 1) * ty.width() + 255) & ~size_t(255);
    }
    char* arena = (char*)dev_alloc(std::max<size_t>(total, 256));
    t->owned->p.push_back(arena);
    for (size_t j = 0; j < types.size(); j++) {
        Column col = out_column(types[j], count);
        col.values = arena + offs[j];
        t->cols.push_back(col);
    }
    return t.release();
}

Launch prepare_args(rf_query* q, const rf_arg* args, int64_t n_args) {
    const Plan& plan = q->plan;
    check_args(plan, args, n_args);
    Launch l;
    l.spec = &get_spec(q, args, n_args);
    const KernelSource& ks = l.spec->ks;
    EvalCtx c{&plan, args, n_args, {}, {}};
    int64_t lo = eval_idx(plan.start, c), end = eval_idx(plan.end, c);
    if (plan.source == Plan::TABLE_ROWS) {
        const rf_table* dt = args[plan.driver].u.table;
        // Rows past the scanned table's end are never read (the reference
        // does not check; reading them is undefined there).
        int64_t n = dt->rows();
        if (end > n) end = n;
        if (lo < 0) lo = 0;
    }
    l.lo = lo;
    l.end = end;
    RfScanArgs& A = l.args;
    std::memset(&A, 0, sizeof A);
    const int64_t items = ks.tune.items;
    const int64_t tile = (int64_t)ks.tune.sub * 64 * items;  // rows per wave-tile
    int64_t base = lo >= 0 ? lo - (lo % items) : -(((-lo) + items - 1) / items) * items;
    A.start = base;
    A.lo = lo;
    A.end = end;
    A.ntiles = end > lo ? (end - base + tile - 1) / tile : 0;
    l.empty = A.ntiles == 0;
    // Columns.
    for (size_t s = 0; s < ks.cols.size(); s++) {
        const ColSlot& cs = ks.cols[s];
        const rf_table* t = args[cs.table].u.table;
        const Column& col = t->cols.at(cs.col);
        A.col[s] = col.values;
        A.valid[s] = (const uint16_t*)col.validity;
    }
    for (size_t j = 0; j < plan.probes.size(); j++) {
        if ((int)j >= RF_MAXPROBES) throw Unsupported("more than 4 probed tables");
        if (plan.probes[j].external) {
            if (j >= q->probe_override.size()) throw IRError("external probe without an index");
            A.probe[j] = q->probe_override[j];
            continue;
        }
        rf_table* t = args[plan.probes[j].table].u.table;
        if (!t->has_index) table_build_index(t);
        A.probe[j] = probe_of(t);
    }
    fill_params(ks, c, A.p);
    return l;
}

int scan_grid(const Launch& l) {
    // Launch width: waves_per_cu resident waves per CU (bounded by occupancy),
    // each owning a contiguous range of wave-tiles.
    const KernelSource& ks = l.spec->ks;
    if (ks.sequential) return 1;  // one wave folds every row in order
    const Module& m = *l.spec->mod;
    const int waves_per_block = ks.tune.block / 64;
    const int64_t blocks_per_cu = std::max<int64_t>(
        1, std::min<int64_t>(m.max_blocks_per_cu, ks.tune.waves_per_cu / waves_per_block));
    int64_t grid = std::min<int64_t>((l.args.ntiles + waves_per_block - 1) / waves_per_block,
                                     (int64_t)device().num_cus * blocks_per_cu);
    return (int)std::max<int64_t>(grid, 1);
}

Launch prepare(rf_query* q, const rf_arg* args, int64_t n_args, bool* /*unused*/) {
    Launch l = prepare_args(q, args, n_args);
    const KernelSource& ks = l.spec->ks;
    RfScanArgs& A = l.args;
    const int64_t tile = (int64_t)ks.tune.sub * 64 * ks.tune.items;
    const int64_t lo = l.lo, end = l.end;
    // Workspace.
    l.grid = scan_grid(l);
    const int64_t nw = (int64_t)l.grid * (ks.tune.block / 64);  // waves in the launch
    size_t nparts = ks.det ? (size_t)std::max<int64_t>(A.ntiles, 1) : (size_t)nw;
    size_t need = std::max<size_t>(nparts * std::max(ks.acc_bytes, 8), 256);
    if (need > q->partials_bytes) {
        dev_free(q->partials);
        q->partials = dev_alloc(need);
        q->partials_bytes = need;
    }
    if (!q->result) q->result = dev_alloc(4096);
    if (!q->counters) {
        q->counters = dev_alloc(256);
        dev_fill_u64(q->counters, 0, 32);
    }
    if (!q->out_count) q->out_count = dev_alloc(256);
    if (!q->host_stage) RF_HIP(hipHostMalloc(&q->host_stage, 8192, hipHostMallocDefault));
    if (ks.has_sel) {
        // Per-wave staging regions cover every wave-tile a wave can own.
        const int64_t stage_rows = ((A.ntiles + nw - 1) / nw) * tile;
        const size_t stage_total = (size_t)std::max<int64_t>(stage_rows * nw, 1);
        const size_t rows = (size_t)std::max<int64_t>(end - lo, 1);
        if ((size_t)nw > q->waves_cap) {
            dev_free(q->wave_arrays);
            q->waves_cap = std::max<size_t>(nw, 1024);
            q->wave_arrays = dev_alloc(q->waves_cap * 2 * 8);
        }
        if (stage_total > q->stage_rows_cap || rows > q->out_rows_cap ||
            q->outs.size() != ks.out_types.size() || q->stages.size() != ks.stage_types.size()) {
            for (void* p : q->outs) dev_free(p);
            for (void* p : q->stages) dev_free(p);
            q->outs.clear();
            q->stages.clear();
            for (auto& t : ks.out_types) q->outs.push_back(dev_alloc(rows * t.width()));
            for (auto& t : ks.stage_types) q->stages.push_back(dev_alloc(stage_total * t.width()));
            q->out_rows_cap = rows;
            q->stage_rows_cap = stage_total;
        }
        for (size_t j = 0; j < q->outs.size(); j++) A.out[j] = q->outs[j];
        A.out_cap = (int64_t)q->out_rows_cap;
        for (size_t j = 0; j < q->stages.size(); j++) A.stage[j] = q->stages[j];
        A.stage_rows = stage_rows;
        A.nwaves = nw;
        A.wave_cnt = (int64_t*)q->wave_arrays;
        A.wave_dst = A.wave_cnt + q->waves_cap;
    }
    A.partials = q->partials;
    A.result = q->result;
    A.counters = (uint32_t*)q->counters;
    A.out_count = (int64_t*)q->out_count;
    return l;
}

void launch(rf_query* /*q*/, Launch& l, hipStream_t s, bool scan_only) {
    if (l.empty) return;
    void* params[] = {&l.args};
    RF_HIP(hipModuleLaunchKernel(l.spec->mod->fn, l.grid, 1, 1, l.spec->ks.tune.block, 1, 1, 0,
                                 s, params, nullptr));
    if (l.spec->ks.has_sel && !scan_only) {
        const int g = (int)std::min<int64_t>(l.args.nwaves, 8192);
        RF_HIP(hipModuleLaunchKernel(l.spec->mod->gather, g, 1, 1, 256, 1, 1, 0, s, params,
                                     nullptr));
    }
}

// The selection placement kernel alone (after a scan-only launch).
static void launch_gather(Launch& l, hipStream_t s) {
    if (l.empty || !l.spec->ks.has_sel) return;
    void* params[] = {&l.args};
    const int g = (int)std::min<int64_t>(l.args.nwaves, 8192);
    RF_HIP(hipModuleLaunchKernel(l.spec->mod->gather, g, 1, 1, 256, 1, 1, 0, s, params, nullptr));
}

static HostVal finish_reduce(const Plan& plan, const KernelSource& ks, size_t comp,
                             const char* acc_bytes, bool empty, EvalCtx& c) {
    const ReduceComp& rc = plan.reduces[comp];
    std::vector<HostVal> fields;
    for (size_t f = 0; f < rc.fields.size(); f++) {
        const Field& fl = rc.fields[f];
        HostVal v = convert(eval(fl.init, c), fl.type);
        if (ks.sequential && !empty) {
            // The kernel folded from the init: its state is the result.
            v = read_prim(fl.type, acc_bytes + ks.acc_offset[f]);
        } else if (fl.updated && !empty) {
            for (size_t a = 0; a < ks.acc.size(); a++) {
                if (ks.acc[a].comp == (int)comp && ks.acc[a].field == (int)f) {
                    HostVal dv = read_prim(fl.type, acc_bytes + ks.acc_offset[a]);
                    v = combine_host(fl.op, v, dv, fl.type);
                }
            }
        }
        fields.push_back(v);
    }
    // Re-nest into the state struct.
    std::function<HostVal(const Type&, size_t&)> nest = [&](const Type& t, size_t& k) {
        if (t.is_struct()) {
            HostVal h;
            h.type = t;
            for (auto& d : t.dtypes) h.fields.push_back(nest(d, k));
            return h;
        }
        return fields.at(k++);
    };
    size_t k = 0;
    return nest(rc.state_type, k);
}

// ---------------------------------------------------------- group join

GroupJoinState::~GroupJoinState() {
    dev_free(small);
    dev_free(keys2);
    dev_free(rows2);
    dev_free(hist);
    dev_free(rank2);
    for (void* p : pay2) dev_free(p);
    dev_free(seg);
    dev_free(part_small);
    dev_free(prank);
    for (void* p : ppay) dev_free(p);
    dev_free(lds_part);
}

GjBuild::~GjBuild() {
    dev_free(keys);
    dev_free(rows);
    dev_free(hkeys);
    dev_free(hrows);
    dev_free(filter);
    dev_free(dense_bits);
    dev_free(dense_cum);
    dev_free(dense_pack);
}

RfProbeDev GjBuild::probe() const {
    RfProbeDev h{};
    if (dense_bits) {
        h.dense_bits = (const uint64_t*)dense_bits;
        h.dense_cum = (const uint32_t*)dense_cum;
        h.dense_pack = (const uint64_t*)dense_pack;
        h.dense_base = kmin;
        h.dense_len = dense_len;
        return h;
    }
    h.keys = (const int64_t*)hkeys;
    h.rows = (const int64_t*)hrows;
    h.mask = index_cap - 1;
    h.min_key_row = (kmin == RF_EMPTY_KEY) ? 0 : -1;  // the smallest key has rank 0
    h.filter = (const uint32_t*)filter;
    h.fmask = filter_bits - 1;
    return h;
}

static void grow(void*& p, size_t& cap, size_t need) {
    if (need <= cap && p) return;
    dev_free(p);
    p = nullptr;
    p = dev_alloc(std::max<size_t>(need, 256));
    cap = need;
}

static std::vector<std::vector<bool>> all_valid(const Plan& plan) {
    std::vector<std::vector<bool>> nl(plan.inputs.size());
    for (size_t i = 0; i < plan.inputs.size(); i++)
        nl[i].assign(plan.inputs[i]->type.dtypes.size(), false);
    return nl;
}

static std::string all_valid_sig(const Plan& plan) {
    std::string sig;
    for (size_t i = 0; i < plan.inputs.size(); i++) {
        if (!plan.inputs[i]->type.is_vector()) continue;
        sig += std::string(plan.inputs[i]->type.dtypes.size(), '0') + "|";
    }
    return sig;
}

static std::unique_ptr<rf_query> sub_query(const Plan& plan, const Tuning& tune) {
    auto q = std::make_unique<rf_query>();
    q->plan = plan;
    q->det = false;
    q->tune = tune;
    Spec sp;
    sp.ks = gen_scan(plan, all_valid(plan), tune, false);
    sp.mod = &jit_module(sp.ks);
    q->specs.emplace(all_valid_sig(plan), std::move(sp));
    return q;
}

void setup_group_join(rf_query* q, const Tuning* probe_tune) {
    const GroupJoinPlan& gp = *q->plan.gj;
    auto G = std::make_unique<GroupJoinState>();
    G->build_q = sub_query(gp.build_plan, default_tuning(gp.build_plan));
    // The probe scan against a rank map (dense build keys); the one against
    // a hash index is compiled on first use (gj_probe_query).
    Plan dp = gp.probe_plan;
    dp.dense_probes = true;
    G->probe_q = sub_query(dp, probe_tune ? *probe_tune : default_tuning(dp));
    G->probe_q->probe_override.resize(gp.probe_plan.probes.size());
    G->probe_tuned = probe_tune != nullptr;
    Spec fs;
    fs.ks = gen_group_fold(gp, all_valid(q->plan));
    fs.mod = &jit_module(fs.ks);
    G->fold_specs.emplace(all_valid_sig(q->plan), std::move(fs));
    q->tune = G->probe_q->tune;
    q->gj = std::move(G);
}

std::vector<KernelSource> group_join_sources(const Plan& plan,
                                             const std::vector<std::vector<bool>>& nullable) {
    const GroupJoinPlan& gp = *plan.gj;
    Plan dp = gp.probe_plan;
    dp.dense_probes = true;
    std::vector<KernelSource> v = {gen_scan(gp.build_plan, nullable, default_tuning(gp.build_plan), false),
                                   gen_scan(dp, nullable, default_tuning(dp), false),
                                   gen_scan(gp.probe_plan, nullable, default_tuning(gp.probe_plan), false),
                                   gen_group_fold(gp, nullable)};
    if (gj_agg_ok(gp)) {
        Plan lp = dp;
        lp.lds_agg = &gp;
        v.push_back(gen_scan(lp, nullable, default_tuning(dp), false));
    }
    return v;
}

Spec& gj_fold_spec(rf_query* q, const rf_arg* args, int64_t n_args) {
    std::string sig = null_signature(q->plan.inputs, args, n_args);
    auto& specs = q->gj->fold_specs;
    auto it = specs.find(sig);
    if (it != specs.end()) return it->second;
    std::vector<std::vector<bool>> nl(q->plan.inputs.size());
    for (size_t i = 0; i < q->plan.inputs.size(); i++) {
        if (!q->plan.inputs[i]->type.is_vector()) continue;
        for (auto& c : args[i].u.table->cols) nl[i].push_back(c.validity != nullptr);
    }
    Spec sp;
    sp.ks = gen_group_fold(*q->plan.gj, nl);
    sp.mod = &jit_module(sp.ks);
    return specs.emplace(sig, std::move(sp)).first->second;
}

struct GjTiming {
    hipEvent_t scan0 = nullptr, scan1 = nullptr;  // around the probe scan kernel
};

int bits_of(uint64_t range) {
    int b = 0;
    while (b < 64 && (range >> b) != 0) b++;
    return b;
}

// Builds the resident build side (two host synchronisations: the group
// count and key range, then the duplicate-key check).
static const char* kDupMsg =
    "group join: the build table holds a qualifying key more than once "
    "(each group key must occur in one build row), or its key column changed "
    "after the table's key range was taken";

static std::shared_ptr<GjBuild> make_build(rf_query* q, const rf_arg* args, int64_t n_args,
                                           bool defer_check) {
    Device& d = device();
    hipStream_t s = d.stream;
    GroupJoinState& G = *q->gj;
    const GroupJoinPlan& gp = *q->plan.gj;
    rf_query* bq = G.build_q.get();
    if (!G.small) G.small = dev_alloc(256);
    uint64_t* small = (uint64_t*)G.small;  // [0..1] key range, [2] duplicate flag, [3] scratch
    auto B = std::make_shared<GjBuild>();

    // 1. Build selection: (key, build row) of the qualifying build rows, and
    // their key range folded in the same scan; the completing block writes
    // the count and the range straight into pinned host memory.
    Launch lb = prepare(bq, args, n_args, nullptr);
    char* hs = (char*)bq->host_stage;  // pinned, allocated by prepare
    if (!bq->host_stage_dev) RF_HIP(hipHostGetDevicePointer(&bq->host_stage_dev, bq->host_stage, 0));

    // Dense keys without a host synchronisation. When the key range of the
    // whole build table (column 0, cached per table) is dense, the rank map
    // covers that range, the selection's count stays in device memory and
    // the rank-map kernels read it there; the count and the duplicate check
    // reach the host at the run's last synchronisation. This saves the build
    // its host round trip: waiting for the selection, then issuing the
    // rank-map launches to an idle GPU (about 35 us of the C4 query with the
    // build).
    {
        rf_table* bt = args[gp.build].u.table;
        const int64_t nb = bt->rows();
        int64_t tmin = 0, tmax = 0;
        if (!lb.empty && nb < ((int64_t)1 << 31) && table_key_range(bt, tmin, tmax)) {
            const uint64_t range = (uint64_t)tmax - (uint64_t)tmin + 1;
            const int64_t words = (int64_t)((range + 63) / 64);
            if (range != 0 && range <= 64 * (uint64_t)nb + ((uint64_t)1 << 20) &&
                words <= kRankFusedWords) {
                // The selection first: the GPU runs it while the host issues
                // the rest.
                launch(bq, lb, s, false);
                B->dense_bits = dev_alloc((size_t)words * 8);
                B->dense_cum = dev_alloc((size_t)aot_rank_cum_entries(words) * 4);
                B->dense_pack = dev_alloc((size_t)words * 16);
                B->dense_len = range;
                B->keys = dev_alloc((size_t)nb * 8);
                B->rows = dev_alloc((size_t)nb * 8);
                dev_fill_u64(B->dense_bits, 0, words);
                const int64_t* kdev = lb.args.out_count;
                char* hd = (char*)bq->host_stage_dev;
                RF_HIP(aot_rank_build_small((const int64_t*)bq->outs[0], nb, kdev, tmin,
                                            B->dense_bits, words, B->dense_cum, B->dense_pack,
                                            hd + 6144, kdev, (int64_t*)(hd + 6160), s));
                RF_HIP(aot_rank_place((const int64_t*)bq->outs[0], (const int64_t*)bq->outs[1], nb,
                                      kdev, tmin, B->dense_bits, words, B->dense_cum,
                                      (int64_t*)B->keys, (int64_t*)B->rows, s));
                B->kmin = tmin;
                B->groups_ub = nb;
                B->groups_dev = kdev;
                B->pending_groups = (const volatile int64_t*)(hs + 6160);
                B->pending_total = (const volatile uint64_t*)(hs + 6144);
                if (!defer_check) {
                    RF_HIP(hipStreamSynchronize(s));
                    gj_verify(q, *B);
                }
                return B;
            }
        }
    }
    dev_fill_u64(small + 2, 0, 2);
    lb.args.result = bq->host_stage_dev;
    lb.args.out_count = (int64_t*)((char*)bq->host_stage_dev + 4096);
    launch(bq, lb, s, false);
    int64_t K = 0, kmin = 0, kmax = 0;
    if (!lb.empty) {
        RF_HIP(hipStreamSynchronize(s));
        K = *(int64_t*)(hs + 4096);
        const KernelSource& bks = lb.spec->ks;
        for (size_t a = 0; a < bks.acc.size(); a++) {
            int64_t v;
            std::memcpy(&v, hs + bks.acc_offset[a], 8);
            if (bks.acc[a].op == AggOp::MIN) kmin = v;
            else if (bks.acc[a].op == AggOp::MAX) kmax = v;
        }
    }
    if (K == 0) return B;
    if (K >= ((int64_t)1 << 31)) throw Unsupported("more than 2^31 groups");

    const uint64_t range = (uint64_t)kmax - (uint64_t)kmin + 1;
    const char* dup_msg = kDupMsg;

    // 2-3 with dense keys: a rank map (a bit per key of the range and a
    // count per 64 bits: 1.5 bits per slot) when the key range is dense --
    // at most 64 slots per key plus 2^20, so it stays smaller than an index
    // and fits the caches. It is built from the unsorted keys; a key's rank
    // (marked keys below it) is its position in key order, so the keys and
    // rows are placed by rank with no sort. Fewer marked keys than rows means
    // a duplicate key.
    if (kmin != RF_EMPTY_KEY && range != 0 && range <= 64 * (uint64_t)K + ((uint64_t)1 << 20) &&
        range < ((uint64_t)1 << 40)) {
        const int64_t words = (int64_t)((range + 63) / 64);
        B->dense_bits = dev_alloc((size_t)words * 8);
        B->dense_cum = dev_alloc((size_t)aot_rank_cum_entries(words) * 4);
        B->dense_len = range;
        B->keys = dev_alloc((size_t)K * 8);
        B->rows = dev_alloc((size_t)K * 8);
        dev_fill_u64(B->dense_bits, 0, words);
        if (words <= kRankFusedWords) {
            // Marking, one-block counts + packed form, placement by rank:
            // three launches. The count of marked keys goes to pinned host
            // memory; with defer_check the caller reads it after its own
            // synchronisation.
            B->dense_pack = dev_alloc((size_t)words * 16);
            if (!bq->host_stage_dev)
                RF_HIP(hipHostGetDevicePointer(&bq->host_stage_dev, bq->host_stage, 0));
            uint64_t* total_dev = (uint64_t*)((char*)bq->host_stage_dev + 6144);
            RF_HIP(aot_rank_build_small((const int64_t*)bq->outs[0], K, nullptr, kmin,
                                        B->dense_bits, words, B->dense_cum, B->dense_pack,
                                        total_dev, nullptr, nullptr, s));
            RF_HIP(aot_rank_place((const int64_t*)bq->outs[0], (const int64_t*)bq->outs[1], K,
                                  nullptr, kmin, B->dense_bits, words, B->dense_cum,
                                  (int64_t*)B->keys, (int64_t*)B->rows, s));
            B->groups = K;
            B->kmin = kmin;
            B->pending_total = (const volatile uint64_t*)(hs + 6144);
            if (!defer_check) {
                RF_HIP(hipStreamSynchronize(s));
                gj_verify(q, *B);
            }
            return B;
        }
        RF_HIP(aot_rank_build((const int64_t*)bq->outs[0], K, kmin, B->dense_bits, words,
                              B->dense_cum, s));
        RF_HIP(aot_rank_scatter((const int64_t*)bq->outs[0], (const int64_t*)bq->outs[1], K, kmin,
                                B->dense_bits, words, B->dense_cum, (int64_t*)B->keys,
                                (int64_t*)B->rows, small + 2, s));
        B->dense_pack = dev_alloc((size_t)words * 16);
        RF_HIP(aot_rank_pack(B->dense_bits, B->dense_cum, words, B->dense_pack, s));
        RF_HIP(hipMemcpyAsync(hs, small + 2, 8, hipMemcpyDeviceToHost, s));
        RF_HIP(hipStreamSynchronize(s));
        if (*(uint64_t*)hs != (uint64_t)K) throw IRError(dup_msg);
        B->groups = K;
        B->kmin = kmin;
        return B;
    }

    // 2. Sparse keys: sort the build keys (ranks = positions in key order).
    if ((size_t)K * 8 > G.keys_cap || !G.keys2) {
        dev_free(G.keys2);
        dev_free(G.rows2);
        G.keys2 = G.rows2 = nullptr;
        G.keys2 = dev_alloc((size_t)K * 8);
        G.rows2 = dev_alloc((size_t)K * 8);
        G.keys_cap = (size_t)K * 8;
    }
    const size_t hist_need = (size_t)aot_sort_hist_entries(std::max<int64_t>(K, 1)) * 4;
    grow(G.hist, G.hist_cap, hist_need);
    void* kb[2] = {bq->outs[0], G.keys2};
    void* pb[2][8] = {{bq->outs[1]}, {G.rows2}};
    const int32_t pw8[8] = {8};
    int which = 0;
    RF_HIP(aot_radix_sort(kb, pb, pw8, 1, 8, K, bits_of((uint64_t)kmax - (uint64_t)kmin),
                          (uint64_t)kmin, (uint32_t*)G.hist, &which, s));
    B->keys = dev_alloc((size_t)K * 8);
    B->rows = dev_alloc((size_t)K * 8);
    RF_HIP(hipMemcpyAsync(B->keys, kb[which], (size_t)K * 8, hipMemcpyDeviceToDevice, s));
    RF_HIP(hipMemcpyAsync(B->rows, pb[which][0], (size_t)K * 8, hipMemcpyDeviceToDevice, s));
    const int64_t* skeys = (const int64_t*)B->keys;

    // 3. Key -> rank: the open-addressing index and its bitmap filter.
    RF_HIP(aot_adjacent_dups(skeys, K, small + 2, s));
    uint64_t cap = 16;
    while (cap < (uint64_t)G.table_scale * (uint64_t)K) cap <<= 1;
    B->hkeys = dev_alloc(cap * 8);
    B->hrows = dev_alloc(cap * 8);
    dev_fill_u64(B->hkeys, (uint64_t)RF_EMPTY_KEY, (int64_t)cap);
    dev_fill_u64(B->hrows, (uint64_t)INT64_MAX, (int64_t)cap);
    dev_fill_u64(small + 3, (uint64_t)INT64_MAX, 1);
    RF_HIP(aot_index_build(skeys, nullptr, K, B->hkeys, B->hrows, cap - 1, small + 3, s));
    // Bitmap filter: about 32 bits per key, at most 1 MiB (cache-resident).
    uint64_t fbits = 1024;
    while (fbits < 32 * (uint64_t)K && fbits < ((uint64_t)1 << 23)) fbits <<= 1;
    B->filter = dev_alloc(fbits / 8);
    dev_fill_u64(B->filter, 0, (int64_t)(fbits / 64));
    RF_HIP(aot_filter_build(skeys, K, B->filter, fbits - 1, s));
    RF_HIP(hipMemcpyAsync(hs, small + 2, 8, hipMemcpyDeviceToHost, s));
    RF_HIP(hipStreamSynchronize(s));
    if (*(uint64_t*)hs != 0) throw IRError(dup_msg);
    B->groups = K;
    B->kmin = kmin;
    B->index_cap = cap;
    B->filter_bits = fbits;
    return B;
}

void gj_verify(rf_query* q, GjBuild& b) {
    if (b.pending_groups) {
        b.groups = *b.pending_groups;
        b.pending_groups = nullptr;
        b.groups_dev = nullptr;
    }
    if (!b.pending_total) return;
    const uint64_t total = *b.pending_total;
    b.pending_total = nullptr;
    if (total != (uint64_t)b.groups) {
        // The caller holds q->mu (gj_clear_cache would take it again).
        q->gj->builds.clear();
        q->gj->build_order.clear();
        throw IRError(kDupMsg);
    }
}

std::shared_ptr<GjBuild> gj_build(rf_query* q, const rf_arg* args, int64_t n_args,
                                  bool defer_check) {
    GroupJoinState& G = *q->gj;
    const GroupJoinPlan& gp = *q->plan.gj;
    // The build side depends only on the build table and the build plan's
    // scalar arguments.
    std::string bkey = std::to_string(args[gp.build].u.table->uid);
    {
        const Launch lbk = prepare_args(G.build_q.get(), args, n_args);
        const KernelSource& bks = lbk.spec->ks;
        for (size_t i = 0; i < bks.params.size(); i++)
            if (bks.params[i].kind != ParamSlot::CONST) {
                bkey += "|";
                bkey += std::to_string(lbk.args.p[i].i);
            }
    }
    auto it = G.builds.find(bkey);
    if (it != G.builds.end()) {
        if (it->second->pending_total || it->second->pending_groups) {  // a run that built it ended early
            RF_HIP(hipStreamSynchronize(device().stream));
            std::shared_ptr<GjBuild> B = it->second;
            gj_verify(q, *B);
        }
        return it->second;
    }
    std::shared_ptr<GjBuild> B = make_build(q, args, n_args, defer_check);
    if (G.build_order.size() >= GroupJoinState::kBuilds) {
        G.builds.erase(G.build_order.front());
        G.build_order.erase(G.build_order.begin());
    }
    G.builds.emplace(bkey, B);
    G.build_order.push_back(bkey);
    return B;
}

// The kernel that reads a rank map is only ever paired with a build side
// that has one (its probe has no hash-index path).
rf_query* gj_probe_query(rf_query* q, const GjBuild& b) {
    GroupJoinState& G = *q->gj;
    if (b.dense_bits) return G.probe_q.get();
    if (!G.probe_q_hash) {
        // Dependent probes of a hash index: latency bound; more resident
        // waves keep more in flight (measured on the 60M-row probe with
        // sparse keys: 0.82 ms at 16 waves per CU, 0.99 ms at 8).
        const Plan& pp = q->plan.gj->probe_plan;
        Tuning t = G.probe_tuned ? G.probe_q->tune : default_tuning(pp);
        if (!G.probe_tuned) t.waves_per_cu = 16;
        G.probe_q_hash = sub_query(pp, t);
        G.probe_q_hash->probe_override.resize(pp.probes.size());
    }
    return G.probe_q_hash.get();
}

void gj_clear_cache(rf_query* q) {
    std::lock_guard<std::mutex> lk(q->mu);
    if (!q->gj) return;
    q->gj->builds.clear();
    q->gj->build_order.clear();
}

void gj_fold_args(rf_query* q, const Spec& fs, const GjBuild& b, const rf_arg* args,
                  int64_t n_args, RfFoldArgs& A) {
    std::memset(&A, 0, sizeof A);
    A.groups = b.groups;
    A.keys = (const int64_t*)b.keys;
    A.brow = (const int64_t*)b.rows;
    for (size_t c = 0; c < fs.ks.cols.size(); c++) {
        const ColSlot& cs = fs.ks.cols[c];
        const Column& col = args[cs.table].u.table->cols.at(cs.col);
        A.col[c] = col.values;
        A.valid[c] = (const uint16_t*)col.validity;
    }
    EvalCtx c{&q->plan, args, n_args, {}, {}};
    fill_params(fs.ks, c, A.p);
}

// The partitioned fold's layout: 2^shift groups per partition, nparts
// partitions (sized by the group count's bound while it is on the device).
struct PartLayout {
    int shift = 0;
    int64_t nparts = 0;
};

// False when the groups do not suit the partitioned fold (the caller sorts).
static bool part_layout(const GroupJoinPlan& gp, const GjBuild& B, PartLayout& lay) {
    // A build whose key count is still on the device is sized by its bound
    // (the build table's rows); the fold reads the count itself.
    const bool pend = B.pending_groups != nullptr;
    const int64_t K = pend ? B.groups_ub : B.groups;
    int nslot = 0;
    for (auto& f : gp.fields) nslot += f.updated ? 1 : 0;
    // 2^shift groups per partition: about 512 partitions or more (a block
    // each), at least 16 groups per partition (LDS atomics on one slot
    // serialise), at most 64 KiB of LDS slots and 4096 partitions. Under a
    // bound, partitions of at most 512 groups: the partitions past the
    // actual count are empty, and the ones below keep the fold's width.
    int max_shift = 0;
    while (((size_t)std::max(nslot, 1) * 8) << (max_shift + 1) <= 65536) max_shift++;
    int shift = 4;
    while (shift < max_shift && (!pend || shift < 9) && (K >> (shift + 1)) >= 512) shift++;
    while (shift < max_shift && ((K + ((int64_t)1 << shift) - 1) >> shift) > 4096) shift++;
    const int64_t nparts = std::max<int64_t>((K + ((int64_t)1 << shift) - 1) >> shift, 1);
    const int np = (int)gp.pay_types.size();
    size_t row = 4;
    for (auto& t : gp.pay_types) row += (size_t)t.width();
    // rf_part_scatter's LDS: 16 bytes per partition and a 4096- or 8192-row chunk.
    if (nparts > 4096 || np > 8 || K >= ((int64_t)1 << 31) ||
        (size_t)nparts * 16 + 4096 * row > 160 * 1024)
        return false;
    lay.shift = shift;
    lay.nparts = nparts;
    return true;
}

// The default mode's group fold after the probe scan (no gather, no sort):
// the staged joined rows are moved into partitions of 2^shift consecutive
// ranks (rf_part_hist unless the probe scan counted them, rf_part_scatter)
// and a block per partition folds its groups with LDS atomics and writes
// their outputs (rf_gj_agg).
static rf_table* group_agg(rf_query* q, const Spec& fs, const GjBuild& B, const rf_arg* args,
                           int64_t n_args, rf_query* pq, Launch& lp, const PartLayout& lay,
                           bool counted) {
    Device& d = device();
    hipStream_t s = d.stream;
    GroupJoinState& G = *q->gj;
    const GroupJoinPlan& gp = *q->plan.gj;
    const bool pend = B.pending_groups != nullptr;
    const int64_t K = pend ? B.groups_ub : B.groups;
    int nslot = 0;
    for (auto& f : gp.fields) nslot += f.updated ? 1 : 0;
    const int shift = lay.shift;
    const int64_t nparts = lay.nparts;
    const int np = (int)gp.pay_types.size();
    size_t row = 4;  // bytes of a joined row: rank and payload
    for (auto& t : gp.pay_types) row += (size_t)t.width();
    const RfScanArgs& A = lp.args;
    const size_t cap = lp.empty ? 1 : (size_t)std::max<int64_t>(A.stage_rows * A.nwaves, 1);
    const int64_t wpb = 4;  // wave regions per partition block
    if (!lp.empty)  // (the caller grew it before a probe scan that counts)
        grow(G.hist, G.hist_cap, (size_t)aot_part_entries(nparts, A.nwaves, wpb) * 4);
    if (!G.part_small) G.part_small = dev_alloc(256);
    if (cap > G.part_cap || G.ppay.size() != (size_t)np) {
        dev_free(G.prank);
        G.prank = nullptr;
        for (void* p : G.ppay) dev_free(p);
        G.ppay.clear();
        G.prank = dev_alloc(cap * 4);
        for (int j = 0; j < np; j++) G.ppay.push_back(dev_alloc(cap * gp.pay_types[j].width()));
        G.part_cap = cap;
    }
    // Partition starts at pstart[p * stride]: the starts rf_part_scatter
    // publishes, or one zero for every partition when the probe scan ran no
    // rows.
    const uint32_t* pstart = (const uint32_t*)G.part_small;
    int64_t stride = 0;
    if (!lp.empty) {
        const int64_t nb = (A.nwaves + wpb - 1) / wpb;
        pstart = (const uint32_t*)G.hist + nb * nparts + nparts;  // aot_partition's layout
        stride = 1;
    } else {
        dev_fill_u64(G.part_small, 0, 1);
    }
    if (!lp.empty) {
        RfPartArgs P{};
        P.rank = (const uint32_t*)pq->stages[0];
        for (int j = 0; j < np; j++) {
            P.pay[j] = pq->stages[1 + j];
            P.width[j] = gp.pay_types[j].width();
            P.ppay[j] = G.ppay[j];
        }
        P.np = np;
        P.shift = shift;
        P.nparts = nparts;
        P.stage_rows = A.stage_rows;
        P.nwaves = A.nwaves;
        P.waves_per_block = wpb;
        // 8192-row LDS chunks when they fit (runs of a partition twice as
        // long per chunk: 0.739 against 0.743 ms on the join workload).
        P.chunk = (size_t)nparts * 16 + 8192 * row <= 160 * 1024 ? 8192 : 4096;
        P.wave_cnt = A.wave_cnt;
        P.boff = (uint32_t*)G.hist;
        P.prank = (uint32_t*)G.prank;
        P.groups_dev = pend ? B.groups_dev : nullptr;
        P.counted = counted ? 1 : 0;
        RF_HIP(aot_partition(P, s));
    }
    std::unique_ptr<rf_table> out(alloc_table(gp.out_types, K, 1));
    RfFoldArgs F;
    gj_fold_args(q, fs, B, args, n_args, F);
    F.groups = K;
    for (int j = 0; j < np; j++) F.pay[j] = G.ppay[j];
    for (size_t j = 0; j < out->cols.size(); j++) F.out[j] = out->cols[j].values;
    RfAggArgs GA{};
    GA.prank = (const uint32_t*)G.prank;
    GA.pstart = pstart;
    GA.stride = stride;
    GA.shift = shift;
    GA.groups_dev = pend ? B.groups_dev : nullptr;
    void* params[] = {&F, &GA};
    const size_t lds = (size_t)std::max(nslot, 1) * 8 << shift;
    RF_HIP(hipModuleLaunchKernel(fs.mod->agg, (unsigned)nparts, 1, 1, 256, 1, 1, (unsigned)lds, s,
                                 params, nullptr));
    RF_HIP(hipStreamSynchronize(s));
    return out.release();
}

// LDS bytes a probe block may give its pre-aggregation slots (with the
// scan's own LDS, within the 64 KiB a block gets without opting in).
static constexpr int64_t kLdsAggBytes = 48 * 1024;

// Updated fields of a group join's state (LDS slots per group).
static int gj_slots(const GroupJoinPlan& gp) {
    int n = 0;
    for (auto& f : gp.fields) n += f.updated ? 1 : 0;
    return n;
}

// The default mode's group fold for few groups (at most kLdsAggBytes of
// slots per block): the probe scan folds every joined row
// into its block's LDS slots with LDS atomics and writes the slots once per
// block; rf_gj_lds_done combines the blocks' slots per group and writes the
// outputs. No joined row is staged, partitioned or sorted. Float sums are
// in the atomics' order (the default mode's 1e-12 bar).
static rf_table* group_lds_agg(rf_query* q, const Spec& fs, const GjBuild& B, const rf_arg* args,
                               int64_t n_args, int64_t Kb, GjTiming* tm) {
    Device& d = device();
    hipStream_t s = d.stream;
    GroupJoinState& G = *q->gj;
    const GroupJoinPlan& gp = *q->plan.gj;
    const int nslot = gj_slots(gp);
    // The kernel that reads a rank map only ever meets a rank map (as the
    // staging probe scans, gj_probe_query).
    std::unique_ptr<rf_query>& slot = B.dense_bits ? G.probe_q_lds : G.probe_q_lds_hash;
    if (!slot) {
        Plan lp = gp.probe_plan;
        lp.dense_probes = B.dense_bits != nullptr;
        lp.lds_agg = q->plan.gj.get();
        slot = sub_query(lp, gj_probe_query(q, B)->tune);
        slot->probe_override.resize(gp.probe_plan.probes.size());
    }
    rf_query* pl = slot.get();
    const RfProbeDev h = B.probe();
    for (size_t j = 0; j < gp.probe_plan.probes.size(); j++)
        if (gp.probe_plan.probes[j].external) pl->probe_override[j] = h;
    Launch lp = prepare(pl, args, n_args, nullptr);
    const size_t lds = (size_t)nslot * (size_t)Kb * 8;
    if (!lp.empty) {
        grow(G.lds_part, G.lds_part_cap, (size_t)lp.grid * lds);
        lp.args.agg_gpp = Kb;
        lp.args.agg_part = (unsigned long long*)G.lds_part;
        void* params[] = {&lp.args};
        if (tm) RF_HIP(hipEventRecord(tm->scan0, s));
        RF_HIP(hipModuleLaunchKernel(lp.spec->mod->fn, lp.grid, 1, 1, lp.spec->ks.tune.block, 1, 1,
                                     (unsigned)lds, s, params, nullptr));
        if (tm) RF_HIP(hipEventRecord(tm->scan1, s));
    } else if (tm) {
        RF_HIP(hipEventRecord(tm->scan0, s));
        RF_HIP(hipEventRecord(tm->scan1, s));
    }
    std::unique_ptr<rf_table> out(alloc_table(gp.out_types, Kb, 1));
    RfFoldArgs F;
    gj_fold_args(q, fs, B, args, n_args, F);
    F.groups = Kb;
    for (size_t j = 0; j < out->cols.size(); j++) F.out[j] = out->cols[j].values;
    RfAggArgs GA{};
    GA.groups_dev = B.pending_groups ? B.groups_dev : nullptr;
    GA.bpart = (const unsigned long long*)G.lds_part;
    GA.nblk = lp.empty ? 0 : lp.grid;
    GA.gpp = Kb;
    void* params[] = {&F, &GA};
    const int W = gj_lds_done_waves(nslot);
    RF_HIP(hipModuleLaunchKernel(fs.mod->lds_done, (unsigned)((Kb + 15) / 16), 1, 1, W * 64, 1, 1, 0,
                                 s, params, nullptr));
    RF_HIP(hipStreamSynchronize(s));
    return out.release();
}

// Runs a group join (see GroupJoinPlan): the build side from the cache (or
// built), then the probe scan, a stable sort of the joined rows by group
// rank, segment bounds and the per-group fold. Two host synchronisations
// once the build side is resident: the joined row count and the end.
static rf_table* run_group_join(rf_query* q, const rf_arg* args, int64_t n_args,
                                GjTiming* tm) {
    Device& d = device();
    hipStream_t s = d.stream;
    GroupJoinState& G = *q->gj;
    const GroupJoinPlan& gp = *q->plan.gj;
    check_args(q->plan, args, n_args);
    // A new dense build's duplicate-key check is read after this run's last
    // synchronisation (no synchronisation between the build and the probe).
    const std::shared_ptr<GjBuild> B = gj_build(q, args, n_args, true);
    if (!B->pending_groups && B->groups == 0) return alloc_table(gp.out_types, 0, 1);
    auto verified = [&](rf_table* t) {
        try {
            gj_verify(q, *B);
        } catch (...) {
            table_free(t);
            throw;
        }
        // A table sized by a pending count's bound gets the count's rows.
        for (Column& c : t->cols) c.length = c.logical_length = B->groups;
        return t;
    };
    const Spec& fs = gj_fold_spec(q, args, n_args);
    {
        // Few groups (by the count or its bound): pre-aggregation in LDS.
        const int64_t Kb = B->pending_groups ? B->groups_ub : B->groups;
        const int nslot = gj_slots(gp);
        if (!q->det && fs.mod->lds_done && nslot > 0 && Kb > 0 && Kb * nslot * 8 <= kLdsAggBytes)
            return verified(group_lds_agg(q, fs, *B, args, n_args, Kb, tm));
    }
    rf_query* pq = gj_probe_query(q, *B);
    const RfProbeDev h = B->probe();
    for (size_t j = 0; j < gp.probe_plan.probes.size(); j++)
        if (gp.probe_plan.probes[j].external) pq->probe_override[j] = h;

    // 4. Probe scan: (rank, payload) of joined probe rows, in row order. For
    // the default mode's partitioned fold, a probe scan that compacts in LDS
    // also counts its joined rows per partition (blocks of four waves are the
    // partition kernels' blocks), so no counting pass follows.
    Launch lp = prepare(pq, args, n_args, nullptr);
    char* hs = (char*)pq->host_stage;
    PartLayout lay;
    const bool part = !q->det && fs.mod->agg &&
                      lp.spec->ks.stage_types == lp.spec->ks.out_types && part_layout(gp, *B, lay);
    bool counted = false;
    size_t dyn_lds = 0;
    if (part && !lp.empty && lp.spec->ks.lds_emit && lp.spec->ks.tune.block == 256) {
        grow(G.hist, G.hist_cap, (size_t)aot_part_entries(lay.nparts, lp.args.nwaves, 4) * 4);
        lp.args.part_boff = (uint32_t*)G.hist;
        lp.args.part_nparts = lay.nparts;
        lp.args.part_shift = lay.shift;
        lp.args.part_groups_dev = B->pending_groups ? B->groups_dev : nullptr;
        dyn_lds = (size_t)lay.nparts * 4;
        counted = true;
    }
    if (tm) RF_HIP(hipEventRecord(tm->scan0, s));
    if (!lp.empty) {
        void* params[] = {&lp.args};
        RF_HIP(hipModuleLaunchKernel(lp.spec->mod->fn, lp.grid, 1, 1, lp.spec->ks.tune.block, 1, 1,
                                     (unsigned)dyn_lds, s, params, nullptr));
    }
    if (tm) RF_HIP(hipEventRecord(tm->scan1, s));
    if (part) return verified(group_agg(q, fs, *B, args, n_args, pq, lp, lay, counted));
    launch_gather(lp, s);
    RF_HIP(hipMemcpyAsync(hs, pq->out_count, 8, hipMemcpyDeviceToHost, s));
    RF_HIP(hipStreamSynchronize(s));
    gj_verify(q, *B);
    const int64_t K = B->groups;
    if (K == 0) return alloc_table(gp.out_types, 0, 1);
    const int64_t J = lp.empty ? 0 : *(int64_t*)hs;

    // 5. Stable sort of the joined rows by rank.
    const int np = (int)gp.pay_types.size();
    if (J > 0) {
        if ((size_t)J > G.join_cap || G.pay2.size() != (size_t)np) {
            dev_free(G.rank2);
            G.rank2 = nullptr;
            for (void* p : G.pay2) dev_free(p);
            G.pay2.clear();
            G.rank2 = dev_alloc((size_t)J * 4);
            for (int j = 0; j < np; j++) G.pay2.push_back(dev_alloc((size_t)J * gp.pay_types[j].width()));
            G.join_cap = (size_t)J;
        }
        grow(G.hist, G.hist_cap, (size_t)aot_sort_hist_entries(J) * 4);
    }
    void* kj[2] = {pq->outs[0], G.rank2};
    void* pj[2][8] = {};
    int32_t pwj[8] = {};
    for (int j = 0; j < np; j++) {
        pj[0][j] = pq->outs[1 + j];
        pj[1][j] = (size_t)j < G.pay2.size() ? G.pay2[j] : nullptr;  // allocated once J > 0
        pwj[j] = gp.pay_types[j].width();
    }
    int wj = 0;
    if (J > 1)
        RF_HIP(aot_radix_sort(kj, pj, pwj, np, 4, J, bits_of((uint64_t)(K - 1)), 0,
                              (uint32_t*)G.hist, &wj, s));

    // 6. Segment bounds per group.
    grow(G.seg, G.seg_cap, (size_t)K * 16);
    int64_t* lo = (int64_t*)G.seg;
    int64_t* hi = lo + K;
    dev_fill_u64(lo, 0, 2 * K);
    if (J > 0) RF_HIP(aot_seg_bounds((const uint32_t*)kj[wj], J, lo, hi, s));

    // 7. Fold each group and evaluate the outputs.
    std::unique_ptr<rf_table> out(alloc_table(gp.out_types, K, 1));
    RfFoldArgs A;
    gj_fold_args(q, fs, *B, args, n_args, A);
    A.seg_lo = lo;
    A.seg_hi = hi;
    for (int j = 0; j < np; j++) A.pay[j] = pj[wj][j];
    for (size_t j = 0; j < out->cols.size(); j++) A.out[j] = out->cols[j].values;
    void* params[] = {&A};
    if (J >= kWaveFoldRows * K) {  // large groups: a wave per group
        const int g = (int)std::min<int64_t>((K + 3) / 4, 65536);
        RF_HIP(hipModuleLaunchKernel(fs.mod->fold_w, g, 1, 1, 256, 1, 1, 0, s, params, nullptr));
    } else {
        const int g = (int)std::min<int64_t>((K + 255) / 256, 16384);
        RF_HIP(hipModuleLaunchKernel(fs.mod->fn, g, 1, 1, 256, 1, 1, 0, s, params, nullptr));
    }
    RF_HIP(hipStreamSynchronize(s));
    return out.release();
}

// Launches a run reduction (SegReducePlan) on `s`: one row of `out` per run
// of T's run-end column 0.
static void launch_seg(rf_query* q, const SegReducePlan& sp, Spec& spec, const rf_arg* args,
                       int64_t n_args, rf_table* out, hipStream_t s) {
    const Column& c0 = args[sp.table].u.table->cols.at(0);
    RfSegArgs A;
    std::memset(&A, 0, sizeof A);
    if (sp.ranged) {
        // One fold per key in [lb, ub): out holds that many rows.
        EvalCtx c{&q->plan, args, n_args, {}, {}};
        A.runs = out->rows();
        A.key0 = eval_idx(sp.lb, c);
        A.rows = args[sp.table].u.table->rows();
    } else {
        A.runs = c0.length;
        A.rows = c0.logical_length;
        A.run_ends = (const int32_t*)c0.run_ends;
        A.keys = c0.values;
    }
    const int64_t runs = A.runs;
    if (runs == 0) return;
    for (size_t k = 0; k < spec.ks.cols.size(); k++) {
        const ColSlot& cs = spec.ks.cols[k];
        const Column& col = args[cs.table].u.table->cols.at(cs.col);
        A.col[k] = col.values;
        A.valid[k] = (const uint16_t*)col.validity;
    }
    for (size_t j = 0; j < sp.sub.probes.size(); j++) {
        if ((int)j >= RF_MAXPROBES) throw Unsupported("more than 4 probed tables");
        rf_table* pt = args[sp.sub.probes[j].table].u.table;
        if (!pt->has_index) table_build_index(pt);
        A.probe[j] = probe_of(pt);
    }
    for (size_t j = 0; j < out->cols.size(); j++) A.out[j] = out->cols[j].values;
    EvalCtx c{&q->plan, args, n_args, {}, {}};
    fill_params(spec.ks, c, A.p);
    void* params[] = {&A};
    const int g = (int)std::min<int64_t>((runs + 255) / 256, 16384);
    // The wide copy when a block's 256 runs would not fit the small one on
    // average (with a quarter to spare).
    hipFunction_t fn = spec.mod->fn;
    if (spec.mod->seg_wave && A.rows >= 32 * runs) {
        // 32 rows per run or window and more: a wave per fold.
        const int gw = (int)std::min<int64_t>((runs + 3) / 4, 65536);
        RF_HIP(hipModuleLaunchKernel(spec.mod->seg_wave, gw, 1, 1, 256, 1, 1, 0, s, params, nullptr));
        return;
    }
    if (spec.mod->seg_long && !sp.ranged && 5 * 256 * A.rows > 4 * (int64_t)spec.ks.cap * runs)
        fn = spec.mod->seg_long;
    RF_HIP(hipModuleLaunchKernel(fn, g, 1, 1, 256, 1, 1, 0, s, params, nullptr));
}

// A new table of one row per run of T, or per key of a ranged fold (no
// launch).
static rf_table* alloc_seg_table(const SegReducePlan& sp, const rf_arg* args,
                                 rf_query* q = nullptr, int64_t n_args = 0) {
    if (sp.ranged) {
        EvalCtx c{&q->plan, args, n_args, {}, {}};
        const int64_t lb = eval_idx(sp.lb, c), ub = eval_idx(sp.ub, c);
        return alloc_table(sp.out_types, ub > lb ? ub - lb : 0, 1);
    }
    const Column& c0 = args[sp.table].u.table->cols.at(0);
    if (c0.enc != Enc::RUNEND) throw ArrowError("column 0 is not run-end encoded");
    if (c0.validity) throw Unsupported("null run values");
    return alloc_table(sp.out_types, c0.length, 1);
}

// Runs a run reduction (SegReducePlan): one output row per run.
static rf_table* run_seg_reduce(rf_query* q, const rf_arg* args, int64_t n_args) {
    Device& d = device();
    check_args(q->plan, args, n_args);
    Spec& spec = get_spec(q, args, n_args);
    std::unique_ptr<rf_table> out(alloc_seg_table(*q->plan.seg, args, q, n_args));
    launch_seg(q, *q->plan.seg, spec, args, n_args, out.get(), d.stream);
    RF_HIP(hipStreamSynchronize(d.stream));
    return out.release();
}

// The runs table of a run stage (key, every reduction's fields per run),
// launched on the library stream (the scan of it follows on that stream).
static rf_table* run_stage_table(rf_query* q, const rf_arg* args, int64_t n_args) {
    check_args(q->plan, args, n_args);
    Spec& spec = get_run_spec(q, args, n_args);
    const SegReducePlan& sp = q->plan.runs->seg;
    std::unique_ptr<rf_table> out(alloc_seg_table(sp, args));
    // The runs table is internal: bool fields keep their type (one byte
    // each, 0 or 1) where returned tables store them as int8.
    for (size_t j = 0; j < out->cols.size(); j++)
        if (sp.out_types[j].bt == BT::BOOL) out->cols[j].type = types::BOOL_T;
    launch_seg(q, sp, spec, args, n_args, out.get(), device().stream);
    return out.release();
}

// Runs a union merge (Plan::MERGE): merge-path splits per tile, the count
// kernel (which also checks both key columns strictly ascending) and a scan
// of its per-tile counts give each tile's first output row and the row count
// (one host synchronisation); the emit kernel writes the rows into a table
// of exactly that many rows.
// ev: when set, events recorded around the emit kernel (timing).
static rf_table* run_merge(rf_query* q, const rf_arg* args, int64_t n_args,
                           hipEvent_t* ev = nullptr) {
    Device& d = device();
    Launch l = prepare_args(q, args, n_args);
    const KernelSource& ks = l.spec->ks;
    const Plan& plan = q->plan;
    const rf_table* ta = args[plan.merge_a].u.table;
    const rf_table* tb = args[plan.merge_b].u.table;
    for (const rf_table* t : {ta, tb}) {
        const Column& k = t->cols.at(0);
        if (k.enc != Enc::FLAT || k.type.bt != BT::INT64)
            throw ArrowError("a union merge needs a flat int64 key column");
    }
    const int64_t na = ta->rows(), nb = tb->rows(), M = na + nb;
    if (M >= (int64_t(1) << 32) - RF_MERGE_TILE)
        throw Unsupported("a union merge of 2^32 rows or more");
    if (M == 0) return alloc_table(ks.out_types, 0, 1);
    const int64_t tiles = (M + RF_MERGE_TILE - 1) / RF_MERGE_TILE;
    const int64_t entries = aot_scan_entries(tiles + 1);
    OwnedBuffers ws;
    const size_t split_bytes = ((size_t)(tiles + 1) * 8 + 255) & ~size_t(255);
    char* w = (char*)dev_alloc(256 + split_bytes + (size_t)entries * 4);
    ws.p.push_back(w);
    uint32_t* flag = (uint32_t*)w;
    int64_t* splits = (int64_t*)(w + 256);
    uint32_t* bcnt = (uint32_t*)(w + 256 + split_bytes);
    const int64_t* ka = (const int64_t*)ta->cols[0].values;
    const int64_t* kb = (const int64_t*)tb->cols[0].values;
    dev_fill_u64(flag, 0, 1);
    RF_HIP(aot_merge_splits(ka, na, kb, nb, plan.merge_off[0], plan.merge_off[1], RF_MERGE_TILE,
                            tiles + 1, splits, d.stream));
    l.args.partials = bcnt;
    l.args.counters = flag;
    l.args.wave_dst = splits;
    void* params[] = {&l.args};
    RF_HIP(hipModuleLaunchKernel(l.spec->mod->fn, (unsigned)(tiles + 1), 1, 1, RF_MERGE_BLOCK, 1,
                                 1, 0, d.stream, params, nullptr));
    RF_HIP(aot_scan_u32(bcnt, tiles + 1, d.stream));
    if (!q->host_stage) RF_HIP(hipHostMalloc(&q->host_stage, 8192, hipHostMallocDefault));
    uint32_t* host = (uint32_t*)q->host_stage;
    RF_HIP(hipMemcpyAsync(host, bcnt + tiles, 4, hipMemcpyDeviceToHost, d.stream));
    RF_HIP(hipMemcpyAsync(host + 1, flag, 4, hipMemcpyDeviceToHost, d.stream));
    RF_HIP(hipStreamSynchronize(d.stream));
    if (host[1] != 0)
        throw IRError("a union merge needs the keys of both vectors strictly ascending"
                      " (after the shift of a shifted side)");
    const int64_t total = host[0];
    std::unique_ptr<rf_table> out(alloc_table(ks.out_types, total, 1));
    if (total > 0) {
        for (size_t k = 0; k < ks.out_types.size(); k++) l.args.out[k] = out->cols[k].values;
        l.args.out_cap = total;
        if (ev) RF_HIP(hipEventRecord(ev[0], d.stream));
        RF_HIP(hipModuleLaunchKernel(l.spec->mod->emit, (unsigned)tiles, 1, 1, RF_MERGE_BLOCK, 1, 1,
                                     0, d.stream, params, nullptr));
        if (ev) RF_HIP(hipEventRecord(ev[1], d.stream));
        RF_HIP(hipStreamSynchronize(d.stream));
    }
    return out.release();
}

// Runs a query's preludes on the caller's arguments: `full` gets those
// arguments and then each prelude's result table (owned by `mats`).
static void run_preludes(rf_query* q, const rf_arg* args, int64_t n_args,
                         std::vector<std::unique_ptr<rf_table>>& mats, std::vector<rf_arg>& full) {
    if (n_args != q->n_user_args)
        throw IRError("function '" + q->plan.func_name + "' takes " +
                      std::to_string(q->n_user_args) + " arguments, got " + std::to_string(n_args));
    full.assign(args, args + n_args);
    for (size_t i = 0; i < q->pre.size(); i++) {
        rf_table* t = nullptr;
        run_query(q->pre[i].get(), args, n_args, nullptr, &t);
        if (!t) throw IRError("a computed vector produced no table");
        mats.emplace_back(t);
        if (q->pre_index[i]) table_build_index(t);
        rf_arg a{};
        a.kind = 0;
        a.u.table = t;
        full.push_back(a);
    }
}

// A composite query (Plan::composite): each component's query on the same
// arguments, then the output over their results. Tables the output does not
// hold are freed.
static void run_composite(rf_query* q, const rf_arg* args, int64_t n_args, void* output_addr,
                          rf_table** out_table) {
    const Plan& plan = q->plan;
    EvalCtx c{&plan, args, n_args, {}, {}};
    for (auto& kv : plan.tbl) c.tbl[kv.first->name] = kv.second;
    std::vector<rf_table*> tables;
    try {
        for (size_t i = 0; i < q->parts.size(); i++) {
            rf_query* part = q->parts[i].get();
            const Type& ty = part->plan.output_type;
            if (ty.is_vector()) {
                rf_table* t = nullptr;
                run_query(part, args, n_args, nullptr, &t);
                tables.push_back(t);
                HostVal hv;
                hv.type = ty;
                hv.table = t;
                c.comps[q->part_names[i]] = hv;
            } else {
                size_t sz = 0, al = 0;
                c_layout(ty, sz, al);
                std::vector<char> buf(std::max<size_t>(sz, 8));
                run_query(part, args, n_args, buf.data(), nullptr);
                c.comps[q->part_names[i]] = read_value(ty, buf.data());
            }
        }
        HostVal out = eval(plan.output, c);
        std::set<void*> kept;
        std::function<void(const HostVal&)> held = [&](const HostVal& v) {
            if (v.table) kept.insert(v.table);
            for (auto& f : v.fields) held(f);
        };
        held(out);
        if (plan.output_type.is_vector()) {
            if (!out_table) throw IRError("an Op output needs out_table");
            *out_table = (rf_table*)out.table;
            if (output_addr) *(void**)output_addr = out.table;
        } else if (output_addr) {
            write_value(out, plan.output_type, (char*)output_addr);
        } else {
            throw IRError("no output address");
        }
        for (rf_table* t : tables)
            if (!kept.count(t)) table_free(t);
    } catch (...) {
        for (rf_table* t : tables) table_free(t);
        throw;
    }
}

void run_query(rf_query* q, const rf_arg* args, int64_t n_args, void* output_addr,
               rf_table** out_table) {
    std::vector<std::unique_ptr<rf_table>> mats;
    std::vector<rf_arg> full;
    if (!q->pre.empty()) {
        run_preludes(q, args, n_args, mats, full);
        args = full.data();
        n_args = (int64_t)full.size();
    }
    if (q->plan.composite) {
        run_composite(q, args, n_args, output_addr, out_table);
        return;
    }
    std::lock_guard<std::mutex> lk(q->mu);
    // A run stage first folds T's runs; the plan then scans that table in
    // T's place (same stream, so no synchronisation in between).
    std::unique_ptr<rf_table> runs;
    std::vector<rf_arg> staged;
    if (q->plan.runs) {
        runs.reset(run_stage_table(q, args, n_args));
        staged.assign(args, args + n_args);
        staged[q->plan.runs->table].u.table = runs.get();
        args = staged.data();
    }
    if (q->plan.seg) {
        if (!out_table) throw IRError("an Op output needs out_table");
        rf_table* t = run_seg_reduce(q, args, n_args);
        *out_table = t;
        if (output_addr) *(void**)output_addr = t;
        return;
    }
    if (q->gj) {
        if (!out_table) throw IRError("an Op output needs out_table");
        rf_table* t = run_group_join(q, args, n_args, nullptr);
        *out_table = t;
        if (output_addr) *(void**)output_addr = t;
        return;
    }
    if (q->plan.source == Plan::MERGE) {
        rf_table* t = run_merge(q, args, n_args);
        finish_output(q, get_spec(q, args, n_args).ks, args, n_args, nullptr, true, t,
                      output_addr, out_table);
        return;
    }
    Device& d = device();
    Launch l = prepare(q, args, n_args, nullptr);
    const KernelSource& ks = l.spec->ks;
    const int64_t dim = q->plan.ops.empty() ? 1 : q->plan.ops[0].n_iters;
    // A selection is placed straight into the returned table, sized from
    // the previous run's row count (twice it, plus 4096 rows); rows that do
    // not fit are placed again into a table of the exact size. One host
    // synchronisation in the common case.
    std::unique_ptr<rf_table> sel_t;
    int64_t sel_cap = 0;
    if (ks.has_sel && !l.empty) {
        sel_cap = std::min<int64_t>(std::max<int64_t>(l.end - l.lo, 1), 2 * q->sel_hint + 4096);
        sel_t.reset(alloc_table(ks.out_types, sel_cap, dim));
        for (size_t j = 0; j < ks.out_types.size(); j++) l.args.out[j] = sel_t->cols[j].values;
        l.args.out_cap = sel_cap;
    }
    // The completing block writes the reduction result and the selection
    // count straight into the query's pinned host stage (mapped into the
    // device's address space): no copy kernel follows the scan, and the
    // values are the host's once the stream is synchronised.
    char* stage = (char*)q->host_stage;
    if (!q->host_stage_dev) RF_HIP(hipHostGetDevicePointer(&q->host_stage_dev, q->host_stage, 0));
    if ((size_t)ks.acc_bytes > 4096) throw Unsupported("reduction state larger than 4 KiB");
    l.args.result = q->host_stage_dev;
    l.args.out_count = (int64_t*)((char*)q->host_stage_dev + 4096);
    launch(q, l, d.stream, false);
    RF_HIP(hipStreamSynchronize(d.stream));
    rf_table* sel = nullptr;
    if (!q->plan.ops.empty()) {
        const int64_t count = l.empty ? 0 : *(int64_t*)(stage + 4096);
        if (!sel_t || count > sel_cap) {
            sel_t.reset(alloc_table(ks.out_types, count, dim));
            if (count > 0) {
                for (size_t j = 0; j < ks.out_types.size(); j++) l.args.out[j] = sel_t->cols[j].values;
                l.args.out_cap = count;
                launch_gather(l, d.stream);
                RF_HIP(hipStreamSynchronize(d.stream));
            }
        }
        for (auto& c : sel_t->cols) c.length = c.logical_length = count;
        q->sel_hint = count;
        sel = sel_t.release();
    }
    finish_output(q, ks, args, n_args, stage, l.empty, sel, output_addr, out_table);
}

void finish_output(rf_query* q, const KernelSource& ks, const rf_arg* args, int64_t n_args,
                   const char* acc, bool empty, rf_table* sel, void* output_addr,
                   rf_table** out_table) {
    const Plan& plan = q->plan;
    EvalCtx c{&plan, args, n_args, {}, {}};
    for (auto& kv : plan.tbl) c.tbl[kv.first->name] = kv.second;
    for (size_t r = 0; r < plan.reduces.size(); r++)
        c.comps[plan.reduces[r].sym_name] = finish_reduce(plan, ks, r, acc, empty, c);
    if (!plan.ops.empty()) {
        HostVal hv;
        hv.type = plan.ops[0].vec_type;
        hv.table = sel;
        c.comps[plan.ops[0].sym_name] = hv;
    }
    // Func output: evaluated over the component results (a struct output
    // that holds an Op result carries the table handle in its field).
    HostVal out = eval(plan.output, c);
    if (plan.output_type.is_vector()) {
        if (!out_table) {
            if (sel) table_free(sel);
            throw IRError("an Op output needs out_table");
        }
        *out_table = (rf_table*)out.table;
        if (output_addr) *(void**)output_addr = out.table;
        return;
    }
    if (output_addr) write_value(out, plan.output_type, (char*)output_addr);
    if (out_table) *out_table = sel;
}

rf_table* table_view(const std::vector<Type>& types, const std::vector<void*>& values,
                     int64_t count, int64_t dim) {
    auto t = std::make_unique<rf_table>();
    t->dim = dim;
    for (size_t j = 0; j < types.size(); j++) {
        Column col = out_column(types[j], count);
        col.values = values[j];
        t->cols.push_back(col);
    }
    return t.release();
}

void time_query(rf_query* q, const rf_arg* args, int64_t n_args, int iters, double* ms_all,
                double* ms_scan) {
    Device& d = device();
    *ms_all = *ms_scan = 0.0;
    if (!q->pre.empty() || q->plan.composite) {
        // Whole runs, preludes included (each run synchronises).
        if (iters <= 0) return;
        hipEvent_t e[2];
        for (auto& x : e) RF_HIP(hipEventCreate(&x));
        std::vector<char> out(4096);
        rf_table* t = nullptr;
        run_query(q, args, n_args, out.data(), &t);
        if (t) table_free(t);
        RF_HIP(hipEventRecord(e[0], d.stream));
        for (int i = 0; i < iters; i++) {
            t = nullptr;
            run_query(q, args, n_args, out.data(), &t);
            if (t) table_free(t);
        }
        RF_HIP(hipEventRecord(e[1], d.stream));
        RF_HIP(hipEventSynchronize(e[1]));
        float a = 0;
        RF_HIP(hipEventElapsedTime(&a, e[0], e[1]));
        for (auto& x : e) (void)hipEventDestroy(x);
        *ms_all = *ms_scan = a / iters;
        return;
    }
    std::lock_guard<std::mutex> lk(q->mu);
    if (q->gj) {
        // Whole group joins (host synchronisations included) and the probe
        // scan kernel of each.
        if (iters <= 0) return;
        GjTiming tm;
        hipEvent_t e[2];
        RF_HIP(hipEventCreate(&tm.scan0));
        RF_HIP(hipEventCreate(&tm.scan1));
        for (auto& x : e) RF_HIP(hipEventCreate(&x));
        table_free(run_group_join(q, args, n_args, nullptr));  // warm
        double scan = 0;
        RF_HIP(hipEventRecord(e[0], d.stream));
        for (int i = 0; i < iters; i++) {
            table_free(run_group_join(q, args, n_args, &tm));
            float ms = 0;
            RF_HIP(hipEventElapsedTime(&ms, tm.scan0, tm.scan1));
            scan += ms;
        }
        RF_HIP(hipEventRecord(e[1], d.stream));
        RF_HIP(hipEventSynchronize(e[1]));
        float a = 0;
        RF_HIP(hipEventElapsedTime(&a, e[0], e[1]));
        for (auto& x : e) (void)hipEventDestroy(x);
        (void)hipEventDestroy(tm.scan0);
        (void)hipEventDestroy(tm.scan1);
        *ms_all = a / iters;
        *ms_scan = scan / iters;
        return;
    }
    std::unique_ptr<rf_table> runs;
    std::vector<rf_arg> staged;
    Spec* run_spec = nullptr;
    const rf_arg* user_args = args;
    if (q->plan.runs) {
        // Timed end to end: the runs table's fold, then the scan of it.
        runs.reset(run_stage_table(q, args, n_args));
        RF_HIP(hipStreamSynchronize(d.stream));
        run_spec = &get_run_spec(q, args, n_args);
        staged.assign(args, args + n_args);
        staged[q->plan.runs->table].u.table = runs.get();
        args = staged.data();
    }
    if (q->plan.source == Plan::MERGE) {
        // Whole runs (each synchronises) and the emit kernel of each.
        if (iters <= 0) return;
        hipEvent_t e[4];
        for (auto& x : e) RF_HIP(hipEventCreate(&x));
        table_free(run_merge(q, args, n_args));
        double emit = 0;
        RF_HIP(hipEventRecord(e[0], d.stream));
        for (int i = 0; i < iters; i++) {
            rf_table* t = run_merge(q, args, n_args, e + 2);
            const bool emitted = t->rows() > 0;  // the events are recorded
            table_free(t);
            float ms = 0;
            if (emitted) RF_HIP(hipEventElapsedTime(&ms, e[2], e[3]));
            emit += ms;
        }
        RF_HIP(hipEventRecord(e[1], d.stream));
        RF_HIP(hipEventSynchronize(e[1]));
        float a = 0;
        RF_HIP(hipEventElapsedTime(&a, e[0], e[1]));
        for (auto& x : e) (void)hipEventDestroy(x);
        *ms_all = a / iters;
        *ms_scan = emit / iters;
        return;
    }
    if (q->plan.seg) {
        // Whole runs (each synchronises).
        if (iters <= 0) return;
        auto run = [&] { return run_seg_reduce(q, args, n_args); };
        hipEvent_t e[2];
        for (auto& x : e) RF_HIP(hipEventCreate(&x));
        table_free(run());
        RF_HIP(hipEventRecord(e[0], d.stream));
        for (int i = 0; i < iters; i++) table_free(run());
        RF_HIP(hipEventRecord(e[1], d.stream));
        RF_HIP(hipEventSynchronize(e[1]));
        float a = 0;
        RF_HIP(hipEventElapsedTime(&a, e[0], e[1]));
        for (auto& x : e) (void)hipEventDestroy(x);
        *ms_all = *ms_scan = a / iters;
        return;
    }
    Launch l = prepare(q, args, n_args, nullptr);
    if (l.empty || iters <= 0) return;
    hipEvent_t e[4];
    for (auto& x : e) RF_HIP(hipEventCreate(&x));
    launch(q, l, d.stream, false);  // warm
    RF_HIP(hipEventRecord(e[0], d.stream));
    for (int i = 0; i < iters; i++) {
        if (run_spec)
            launch_seg(q, q->plan.runs->seg, *run_spec, user_args, n_args, runs.get(), d.stream);
        launch(q, l, d.stream, false);
    }
    RF_HIP(hipEventRecord(e[1], d.stream));
    for (int i = 0; i < iters; i++) {
        if (run_spec)  // a run stage: the fold streams the table
            launch_seg(q, q->plan.runs->seg, *run_spec, user_args, n_args, runs.get(), d.stream);
        else
            launch(q, l, d.stream, true);
    }
    RF_HIP(hipEventRecord(e[2], d.stream));
    RF_HIP(hipEventSynchronize(e[2]));
    float a = 0, b = 0;
    RF_HIP(hipEventElapsedTime(&a, e[0], e[1]));
    RF_HIP(hipEventElapsedTime(&b, e[1], e[2]));
    for (auto& x : e) (void)hipEventDestroy(x);
    *ms_all = a / iters;
    *ms_scan = b / iters;
}

}  // namespace rf
