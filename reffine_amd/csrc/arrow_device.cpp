// Arrow C Device data interface for device tables: import of record batches
// that live in HBM (zero copy) or host memory (upload), and export of device
// tables as ArrowDeviceArray (zero copy). The structs are the published ABI
// (arrow/c/abi.h); the reference exchanges only host ArrowArray/ArrowSchema
// (include/reffine/arrow/abi.h:18-69, src/utils/utils.cpp:11-34).
#include <cstring>
#include <list>
#include <string>

#include "engine.hpp"

namespace rf {

namespace {

bool host_accessible(ArrowDeviceType t) {
    return t == ARROW_DEVICE_CPU || t == ARROW_DEVICE_CUDA_HOST || t == ARROW_DEVICE_ROCM_HOST;
}

// The array, moved into shared ownership: released when the last holder
// (a table that borrows its buffers) drops it.
std::shared_ptr<ArrowArray> hold(ArrowArray* a) {
    auto* h = new ArrowArray(*a);
    a->release = nullptr;  // moved
    return std::shared_ptr<ArrowArray>(h, [](ArrowArray* x) {
        if (x->release) x->release(x);
        delete x;
    });
}

// A host copy of one device ArrowArray node (its buffers over
// [0, offset + length) elements, children recursively), for the upload path.
struct HostMirror {
    std::list<std::vector<uint8_t>> bytes;
    std::list<ArrowArray> nodes;
    std::list<std::vector<const void*>> bufs;
    std::list<std::vector<ArrowArray*>> kids;

    const void* copy(const void* dev, size_t n) {
        bytes.emplace_back(std::max<size_t>(n, 1));
        if (dev && n) RF_HIP(hipMemcpy(bytes.back().data(), dev, n, hipMemcpyDeviceToHost));
        return dev ? bytes.back().data() : nullptr;
    }

    // Element width of a primitive Arrow format ("c", "l", "g" ...).
    static size_t width(const std::string& f) {
        if (f == "c" || f == "C" || f == "b") return 1;
        if (f == "s" || f == "S") return 2;
        if (f == "i" || f == "I" || f == "f") return 4;
        if (f == "l" || f == "L" || f == "g") return 8;
        throw ArrowError("unsupported Arrow format '" + f + "'");
    }

    ArrowArray* mirror(const ArrowSchema* s, const ArrowArray* a) {
        nodes.push_back(*a);
        ArrowArray* m = &nodes.back();
        m->release = nullptr;
        m->private_data = nullptr;
        const std::string f = s->format ? s->format : "";
        const int64_t end = a->offset + a->length;
        // References into the lists stay valid while the recursion below
        // appends to them.
        std::vector<const void*>& mb =
            bufs.emplace_back((size_t)std::max<int64_t>(a->n_buffers, 0), nullptr);
        if (f == "+s" || f == "+r") {
            if (a->n_buffers > 0 && a->buffers[0] && a->null_count != 0)
                throw ArrowError("nulls on a struct or run-end array are not supported");
        } else {
            if (a->n_buffers < 2) throw ArrowError("primitive array without a values buffer");
            mb[0] = copy(a->buffers[0], (size_t)(end + 7) / 8);
            mb[1] = copy(a->buffers[1], (size_t)end * width(f));
        }
        m->buffers = mb.data();
        std::vector<ArrowArray*>& mk = kids.emplace_back();
        for (int64_t i = 0; i < a->n_children; i++) mk.push_back(mirror(s->children[i], a->children[i]));
        m->children = mk.data();
        return m;
    }
};

// Zero copy when every column is flat, its values start 16-byte aligned and
// its validity starts on a 16-bit word (the kernels' load granules).
bool borrowable(const ArrowSchema* s, const ArrowArray* a) {
    for (int64_t i = 0; i < a->n_children; i++) {
        const ArrowArray* c = a->children[i];
        const std::string f = s->children[i]->format ? s->children[i]->format : "";
        if (f.empty() || f[0] == '+' || c->n_buffers < 2) return false;
        const size_t w = HostMirror::width(f);
        const int64_t eo = a->offset + c->offset;
        if (((uintptr_t)c->buffers[1] + (uintptr_t)eo * w) % 16 != 0) return false;
        if (c->buffers[0] && c->null_count != 0) {
            if (eo % 16 != 0 || ((uintptr_t)c->buffers[0] + (uintptr_t)eo / 8) % 2 != 0)
                return false;
        }
    }
    return true;
}

}  // namespace

rf_table* table_import_device(const ArrowSchema* schema, ArrowDeviceArray* arr, int64_t dim) {
    if (!schema || !arr) throw ArrowError("null schema or array");
    // The array is moved in: released here on error or after a copy, or when
    // the table that borrows it is freed.
    std::shared_ptr<ArrowArray> held = hold(&arr->array);
    const ArrowArray* a = held.get();
    if (!schema->format || std::string(schema->format) != "+s")
        throw ArrowError("record batch must be a struct array (format \"+s\")");
    if (schema->n_children != a->n_children) throw ArrowError("schema/array child count mismatch");
    if (host_accessible(arr->device_type)) return table_upload(schema, a, dim);
    if (arr->device_type != ARROW_DEVICE_ROCM)
        throw ArrowError("device type " + std::to_string(arr->device_type) +
                         " is neither ROCm nor host memory");
    Device& d = device();
    if (arr->device_id != d.id)
        throw ArrowError("array on device " + std::to_string(arr->device_id) +
                         ", library on device " + std::to_string(d.id));
    if (arr->sync_event) RF_HIP(hipStreamWaitEvent(d.stream, *(hipEvent_t*)arr->sync_event, 0));
    if (dim < 1 || dim > schema->n_children) throw ArrowError("bad dim");
    if (!borrowable(schema, a)) {
        RF_HIP(hipStreamSynchronize(d.stream));
        HostMirror hm;
        ArrowArray* root = hm.mirror(schema, a);
        return table_upload(schema, root, dim);
    }
    auto t = std::make_unique<rf_table>();
    t->dim = dim;
    t->cols.resize(schema->n_children);
    for (int64_t i = 0; i < schema->n_children; i++) {
        const ArrowArray* c = a->children[i];
        Column& col = t->cols[i];
        col.format = schema->children[i]->format;
        if (!format_type_of(col.format, col.type))
            throw ArrowError("unsupported Arrow format '" + col.format + "'");
        const int64_t eo = a->offset + c->offset;
        if (c->length < a->offset + a->length) throw ArrowError("column shorter than the batch");
        col.length = col.logical_length = a->length;
        col.values = (char*)c->buffers[1] + (size_t)eo * col.type.width();
        if (c->buffers[0] && c->null_count != 0) {
            col.validity = (char*)c->buffers[0] + eo / 8;
            col.null_count = c->null_count;
        }
    }
    t->keep.push_back(std::static_pointer_cast<void>(held));
    RF_HIP(hipStreamSynchronize(d.stream));
    return t.release();
}

namespace {

struct ExpSchema {
    std::string format, name;
    std::vector<ArrowSchema*> kids;
};

struct ExpArray {
    std::vector<const void*> bufs;
    std::vector<ArrowArray*> kids;
    std::shared_ptr<OwnedBuffers> own;
    std::vector<std::shared_ptr<void>> keep;
};

void release_schema(ArrowSchema* s) {
    auto* p = (ExpSchema*)s->private_data;
    for (ArrowSchema* c : p->kids) {
        if (c->release) c->release(c);
        delete c;
    }
    delete p;
    s->release = nullptr;
}

void release_array(ArrowArray* a) {
    auto* p = (ExpArray*)a->private_data;
    for (ArrowArray* c : p->kids) {
        if (c->release) c->release(c);
        delete c;
    }
    delete p;
    a->release = nullptr;
}

void make_schema(ArrowSchema* s, const std::string& format, const std::string& name,
                 bool nullable, std::vector<ArrowSchema*> kids) {
    auto* p = new ExpSchema{format, name, std::move(kids)};
    std::memset(s, 0, sizeof *s);
    s->format = p->format.c_str();
    s->name = p->name.c_str();
    s->flags = nullable ? 2 : 0;  // ARROW_FLAG_NULLABLE
    s->n_children = (int64_t)p->kids.size();
    s->children = p->kids.empty() ? nullptr : p->kids.data();
    s->release = release_schema;
    s->private_data = p;
}

void make_array(ArrowArray* a, const rf_table* t, int64_t length, int64_t null_count,
                std::vector<const void*> bufs, std::vector<ArrowArray*> kids) {
    auto* p = new ExpArray{std::move(bufs), std::move(kids), t->owned, t->keep};
    std::memset(a, 0, sizeof *a);
    a->length = length;
    a->null_count = null_count;
    a->n_buffers = (int64_t)p->bufs.size();
    a->buffers = p->bufs.empty() ? nullptr : p->bufs.data();
    a->n_children = (int64_t)p->kids.size();
    a->children = p->kids.empty() ? nullptr : p->kids.data();
    a->release = release_array;
    a->private_data = p;
}

}  // namespace

void table_export_device(const rf_table* t, ArrowSchema* schema, ArrowDeviceArray* arr) {
    if (!t || !schema || !arr) throw ArrowError("null argument");
    Device& d = device();
    std::vector<ArrowSchema*> sk;
    std::vector<ArrowArray*> ak;
    try {
        for (size_t i = 0; i < t->cols.size(); i++) {
            const Column& c = t->cols[i];
            const std::string name = "c" + std::to_string(i);
            const int64_t nulls = c.validity ? -1 : 0;
            auto* cs = new ArrowSchema();
            auto* ca = new ArrowArray();
            sk.push_back(cs);
            ak.push_back(ca);
            if (c.enc == Enc::RUNEND) {
                auto* rs = new ArrowSchema();
                auto* vs = new ArrowSchema();
                make_schema(rs, "i", "run_ends", false, {});
                make_schema(vs, arrow_format_of(c.type), "values", c.validity != nullptr, {});
                make_schema(cs, "+r", name, false, {rs, vs});
                auto* ra = new ArrowArray();
                auto* va = new ArrowArray();
                make_array(ra, t, c.length, 0, {nullptr, c.run_ends}, {});
                make_array(va, t, c.length, nulls, {c.validity, c.values}, {});
                make_array(ca, t, c.logical_length, 0, {}, {ra, va});
            } else {
                make_schema(cs, c.format, name, c.validity != nullptr, {});
                make_array(ca, t, c.length, nulls, {c.validity, c.values}, {});
            }
        }
    } catch (...) {
        for (auto* x : sk) { if (x->release) x->release(x); delete x; }
        for (auto* x : ak) { if (x->release) x->release(x); delete x; }
        throw;
    }
    const int64_t rows = t->cols.empty() ? 0
                         : (t->cols[0].enc == Enc::RUNEND ? t->cols[0].logical_length
                                                          : t->cols[0].length);
    make_schema(schema, "+s", "", false, sk);
    std::memset(arr, 0, sizeof *arr);
    make_array(&arr->array, t, rows, 0, {nullptr}, ak);
    arr->device_id = d.id;
    arr->device_type = ARROW_DEVICE_ROCM;
    arr->sync_event = nullptr;  // complete: every library call synchronises before returning
}

}  // namespace rf
