// extern "C" entry points declared in include/reffine_hip.h. Every call
// catches exceptions and turns them into a status code plus the thread's
// last-error text; nothing crosses the ABI as an exception or exit().
#include <cstring>
#include <functional>
#include <string>

#include "engine.hpp"

namespace {

thread_local std::string g_last_error;

template <class F>
int guard(F&& f) {
    try {
        g_last_error.clear();
        f();
        return RF_OK;
    } catch (const rf::Unsupported& e) {
        g_last_error = e.what();
        return RF_ERR_UNSUPPORTED;
    } catch (const rf::IRError& e) {
        g_last_error = e.what();
        return RF_ERR_INVALID;
    } catch (const rf::NoMem& e) {
        g_last_error = e.what();
        return RF_ERR_NOMEM;
    } catch (const rf::ArrowError& e) {
        g_last_error = e.what();
        return RF_ERR_ARROW;
    } catch (const rf::HipError& e) {
        g_last_error = e.what();
        return RF_ERR_HIP;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return RF_ERR_INVALID;
    } catch (...) {
        g_last_error = "unknown error";
        return RF_ERR_INVALID;
    }
}

void need(bool c, const char* msg) {
    if (!c) throw rf::IRError(msg);
}

}  // namespace

extern "C" {

int rf_init(int device) {
    return guard([&] { rf::init_device(device); });
}

const char* rf_last_error(void) { return g_last_error.c_str(); }

int rf_device_info(int32_t* num_cus, int64_t* total_bytes, int64_t* free_bytes) {
    return guard([&] {
        rf::Device& d = rf::device();
        size_t fr = 0, tot = 0;
        RF_HIP(hipMemGetInfo(&fr, &tot));
        if (num_cus) *num_cus = d.num_cus;
        if (total_bytes) *total_bytes = (int64_t)tot;
        if (free_bytes) *free_bytes = (int64_t)fr;
    });
}

int rf_sync(void) {
    return guard([&] {
        RF_HIP(hipStreamSynchronize(rf::device().stream));
        RF_HIP(hipDeviceSynchronize());
    });
}

int rf_table_upload(const struct ArrowSchema* schema, const struct ArrowArray* array,
                    int64_t dim, rf_table** out) {
    return guard([&] {
        need(out != nullptr, "out is NULL");
        *out = rf::table_upload(schema, array, dim);
    });
}

int rf_table_import_device(const struct ArrowSchema* schema, struct ArrowDeviceArray* array,
                           int64_t dim, rf_table** out) {
    return guard([&] {
        need(out != nullptr, "out is NULL");
        *out = rf::table_import_device(schema, array, dim);
    });
}

int rf_table_export_device(const rf_table* t, struct ArrowSchema* schema,
                           struct ArrowDeviceArray* array) {
    return guard([&] { rf::table_export_device(t, schema, array); });
}

int rf_table_wrap_device(int64_t dim, int64_t n_cols, int64_t n_rows,
                         const char* const* formats, void* const* values,
                         void* const* validity, rf_table** out) {
    return guard([&] {
        need(out && formats && values, "NULL argument");
        *out = rf::table_wrap(dim, n_cols, n_rows, formats, values, validity);
    });
}

int rf_table_shape(const rf_table* t, int64_t* n_rows, int64_t* n_cols) {
    return guard([&] {
        need(t != nullptr, "table is NULL");
        if (n_rows) *n_rows = t->rows();
        if (n_cols) *n_cols = (int64_t)t->cols.size();
    });
}

int rf_table_column(const rf_table* t, int64_t col, void** values, void** validity,
                    int64_t* length, char* format_out, size_t format_cap) {
    return guard([&] {
        need(t != nullptr, "table is NULL");
        need(col >= 0 && col < (int64_t)t->cols.size(), "column out of range");
        const rf::Column& c = t->cols[col];
        if (values) *values = c.values;
        if (validity) *validity = c.validity;
        if (length) *length = c.length;
        if (format_out && format_cap) {
            std::strncpy(format_out, c.format.c_str(), format_cap - 1);
            format_out[format_cap - 1] = 0;
        }
    });
}

int rf_table_download_column(const rf_table* t, int64_t col, void* values_out,
                             size_t values_cap, uint8_t* validity_out, size_t validity_cap) {
    return guard([&] {
        need(t != nullptr, "table is NULL");
        need(col >= 0 && col < (int64_t)t->cols.size(), "column out of range");
        const rf::Column& c = t->cols[col];
        rf::Device& d = rf::device();
        size_t bytes = (size_t)c.length * c.type.width();
        need(values_cap >= bytes, "values buffer too small");
        if (bytes)
            RF_HIP(hipMemcpyAsync(values_out, c.values, bytes, hipMemcpyDeviceToHost, d.stream));
        if (validity_out) {
            size_t vb = ((size_t)c.length + 7) / 8;
            need(validity_cap >= vb, "validity buffer too small");
            if (c.validity)
                RF_HIP(hipMemcpyAsync(validity_out, c.validity, vb, hipMemcpyDeviceToHost, d.stream));
            else
                std::memset(validity_out, 0xff, vb);
        }
        RF_HIP(hipStreamSynchronize(d.stream));
    });
}

int rf_table_build_index(rf_table* t) {
    return guard([&] {
        need(t != nullptr, "table is NULL");
        rf::table_build_index(t);
    });
}

int rf_table_locate(const rf_table* t, const int64_t* d_keys, int64_t n, int64_t* d_rows_out) {
    return guard([&] {
        need(t != nullptr, "table is NULL");
        need(t->has_index, "table has no index (call rf_table_build_index)");
        rf::Device& d = rf::device();
        RF_HIP(rf::aot_index_lookup(rf::probe_of(t), d_keys, n, d_rows_out, d.stream));
        RF_HIP(hipStreamSynchronize(d.stream));
    });
}

int rf_table_free(rf_table* t) {
    return guard([&] { rf::table_free(t); });
}

static std::unique_ptr<rf_query> make_query(const rf::N& f0, uint32_t flags,
                                            const int32_t* tuning, int depth = 0) {
    // Computed vectors nest (Nbody: the force table reads the pair table);
    // a symbol table that refers back to itself would not end.
    need(depth < 8, "computed vectors nested more than 8 deep");
    std::vector<rf::Prelude> pre;
    rf::N f = rf::split_materialized(f0, pre);
    auto q = std::make_unique<rf_query>();
    std::vector<rf::N> parts;
    try {
        q->plan = rf::lower_func(f, false);
    } catch (const rf::Unsupported&) {
        // Components that cannot share one scan run as queries of their own.
        parts = rf::split_components(f);
        if (parts.size() < 2) throw;
        q->plan = rf::Plan();
        q->plan.composite = true;
        q->plan.func_name = f->name;
        q->plan.inputs.assign(f->args.begin() + 1, f->args.end());
        q->plan.tbl = f->tbl;
        q->plan.output = f->args[0];
        q->plan.output_type = f->args[0]->type;
    }
    q->n_user_args = (int64_t)f0->args.size() - 1;
    for (auto& p : pre) {
        q->pre.push_back(make_query(p.func, flags, nullptr, depth + 1));
        q->pre_index.push_back(p.index);
    }
    if (q->plan.composite) {
        for (auto& pf : parts) {
            q->parts.push_back(make_query(pf, flags, nullptr, depth + 1));
            q->part_names.push_back(pf->args[0]->name);
        }
        q->det = (flags & RF_COMPILE_DETERMINISTIC) != 0;
        return q;
    }
    q->det = (flags & RF_COMPILE_DETERMINISTIC) != 0;
    q->tune = rf::default_tuning(q->plan, q->det);
    if (tuning) {
        rf::Tuning t{tuning[0], tuning[1], tuning[2], tuning[3] > 0 ? tuning[3] : 16,
                     tuning[4] != 0 ? 1 : 0, tuning[5] > 0 ? tuning[5] : 0};
        need(t.lsub == 0 || (t.lsub <= t.sub && t.sub % t.lsub == 0),
             "lsub must divide sub");
        need(t.block >= 64 && t.block <= 1024 && t.block % 64 == 0, "block must be 64..1024, x64");
        need(t.items == 1 || t.items == 2 || t.items == 4 || t.items == 8 || t.items == 16,
             "items must be 1, 2, 4, 8 or 16");
        need(t.sub >= 1 && t.sub <= 64, "sub must be 1..64");
        need(t.waves_per_cu >= 1 && t.waves_per_cu <= 64, "waves_per_cu must be 1..64");
        q->tune = t;
    }
    const int scale = (int)((flags >> 8) & 15u);
    need(scale == 0 || scale >= 2, "the index scale must be 2..15 (0: 2)");
    if (q->plan.gj) {
        rf::setup_group_join(q.get(), tuning ? &q->tune : nullptr);
        if (scale) q->gj->table_scale = scale;
        q->tune.index_scale = q->gj->table_scale;
        return q;
    }
    // Compile the all-valid specialisation now so compile errors surface
    // here, as compile_op's do.
    std::vector<std::vector<bool>> nl(q->plan.inputs.size());
    for (size_t i = 0; i < q->plan.inputs.size(); i++)
        nl[i].assign(q->plan.inputs[i]->type.dtypes.size(), false);
    rf::Spec sp;
    sp.ks = q->plan.seg ? rf::gen_seg_fold(*q->plan.seg, q->plan, nl)
                        : rf::gen_scan(q->plan, nl, q->tune, q->det);
    sp.mod = &rf::jit_module(sp.ks);
    std::string sig;
    for (size_t i = 0; i < q->plan.inputs.size(); i++) {
        if (!q->plan.inputs[i]->type.is_vector()) continue;
        sig += std::string(q->plan.inputs[i]->type.dtypes.size(), '0') + "|";
    }
    // Its probes take any index (the form keyed 'h'; a packed rank map over
    // ascending keys gets its own kernel at its first run).
    if (!q->plan.seg && q->plan.source == rf::Plan::TABLE_ROWS)
        for (size_t j = 0; j < q->plan.probes.size(); j++) sig += 'h';
    q->specs.emplace(sig, std::move(sp));
    if (q->plan.runs) {
        // The run stage's fold over T with its own (run-end) type.
        const rf::RunStagePlan& rs = *q->plan.runs;
        std::vector<std::vector<bool>> rnl = nl;
        rnl[rs.table].assign(rs.input->type.dtypes.size(), false);
        rf::Spec rsp;
        rsp.ks = rf::gen_seg_fold(rs.seg, q->plan, rnl);
        rsp.mod = &rf::jit_module(rsp.ks);
        std::string rsig;
        for (size_t i = 0; i < q->plan.inputs.size(); i++) {
            if (!q->plan.inputs[i]->type.is_vector()) continue;
            const size_t nc = (int)i == rs.table ? rs.input->type.dtypes.size()
                                                 : q->plan.inputs[i]->type.dtypes.size();
            rsig += std::string(nc, '0') + "|";
        }
        rsig += std::string(rs.seg.sub.probes.size(), 'h');  // hash-probed tables
        q->run_specs.emplace(rsig, std::move(rsp));
    }
    return q;
}

static void compile_impl(const char* ir_text, uint32_t flags, const int32_t* tuning,
                         rf_query** out) {
    need(ir_text && out, "NULL argument");
    rf::device();
    *out = make_query(rf::parse_func(ir_text), flags, tuning).release();
}

int rf_compile(const char* ir_text, uint32_t flags, rf_query** out) {
    return guard([&] { compile_impl(ir_text, flags, nullptr, out); });
}

int rf_compile_tuned(const char* ir_text, uint32_t flags, const int32_t* tuning, rf_query** out) {
    return guard([&] {
        need(tuning != nullptr, "tuning is NULL");
        compile_impl(ir_text, flags, tuning, out);
    });
}

int rf_compile_check(const char* ir_text, uint32_t flags, char* buf, size_t cap) {
    return guard([&] {
        need(ir_text != nullptr, "NULL argument");
        std::string desc;
        std::vector<rf::KernelSource> kss;
        // Preludes (computed vectors read as tables) first, then the function.
        std::function<void(const rf::N&, int)> add = [&](const rf::N& f0, int depth) {
            need(depth < 8, "computed vectors nested more than 8 deep");
            std::vector<rf::Prelude> pre;
            rf::N f = rf::split_materialized(f0, pre);
            for (auto& p : pre) {
                add(p.func, depth + 1);
                desc += std::string("  materialized as an input") +
                        (p.index ? " with an index" : "") + ": " + rf::to_text(p.func->args[0]) +
                        "\n";
            }
            rf::Plan plan;
            try {
                plan = rf::lower_func(f, false);
            } catch (const rf::Unsupported&) {
                std::vector<rf::N> parts = rf::split_components(f);
                if (parts.size() < 2) throw;
                desc += "func " + f->name + ": independent components, each a query of its own, "
                        "run one after another\n";
                for (auto& pf : parts) add(pf, depth + 1);
                return;
            }
            rf::Tuning tune = rf::default_tuning(plan, (flags & RF_COMPILE_DETERMINISTIC) != 0);
            std::vector<std::vector<bool>> nl(plan.inputs.size());
            for (size_t i = 0; i < plan.inputs.size(); i++)
                nl[i].assign(plan.inputs[i]->type.dtypes.size(), (flags & 2u) != 0);
            if (plan.runs) {
                std::vector<std::vector<bool>> rnl = nl;
                rnl[plan.runs->table].assign(plan.runs->input->type.dtypes.size(),
                                             (flags & 2u) != 0);
                kss.push_back(rf::gen_seg_fold(plan.runs->seg, plan, rnl));
            }
            if (plan.gj) {
                for (auto& ks : rf::group_join_sources(plan, nl)) kss.push_back(ks);
            } else if (plan.seg) {
                kss.push_back(rf::gen_seg_fold(*plan.seg, plan, nl));
            } else {
                kss.push_back(
                    rf::gen_scan(plan, nl, tune, (flags & RF_COMPILE_DETERMINISTIC) != 0));
            }
            desc += plan.describe();
        };
        add(rf::parse_func(ir_text), 0);
        size_t bytes = 0;
        std::string text;
        for (auto& ks : kss) {
            bytes += rf::jit_compile(ks).size();
            text += "\n--- kernel ---\n" + ks.code;
        }
        if (buf && cap) {
            std::string s = desc + "code object bytes: " + std::to_string(bytes) + text;
            std::strncpy(buf, s.c_str(), cap - 1);
            buf[cap - 1] = 0;
        }
    });
}

int rf_query_output_info(const rf_query* q, int64_t* out_bytes, int32_t* out_is_table) {
    return guard([&] {
        need(q != nullptr, "query is NULL");
        size_t sz = 0, al = 0;
        rf::c_layout(q->plan.output_type, sz, al);
        if (out_bytes) *out_bytes = q->plan.output_type.is_vector() ? 0 : (int64_t)sz;
        if (out_is_table) *out_is_table = q->plan.output_type.is_vector() ? 1 : 0;
    });
}

int rf_query_tuning(const rf_query* q, int32_t* tuning) {
    return guard([&] {
        need(q && tuning, "NULL argument");
        tuning[0] = q->tune.block;
        tuning[1] = q->tune.items;
        tuning[2] = q->tune.sub;
        tuning[3] = q->tune.waves_per_cu;
        tuning[4] = q->tune.pipeline;
        tuning[5] = q->tune.lsub;
    });
}

int rf_query_clear_cache(rf_query* q) {
    return guard([&] {
        need(q, "NULL argument");
        rf::gj_clear_cache(q);
    });
}

int rf_query_describe(const rf_query* q, char* buf, size_t cap) {
    return guard([&] {
        need(q && buf && cap, "NULL argument");
        std::string s = q->plan.describe();
        for (size_t i = 0; i < q->parts.size(); i++)
            s += "component " + q->part_names[i] + ":\n" + q->parts[i]->plan.describe();
        auto tn = q->tune;
        s += "tuning: block " + std::to_string(tn.block) + " items " + std::to_string(tn.items) +
             " sub " + std::to_string(tn.sub) + " waves/CU " + std::to_string(tn.waves_per_cu) +
             (tn.pipeline ? " pipelined" : "") +
             (q->gj ? " index scale " + std::to_string(q->gj->table_scale) : std::string()) +
             (q->det ? " deterministic" : "") + "\n";
        if (!q->specs.empty()) s += "--- kernel ---\n" + q->specs.begin()->second.ks.code;
        std::strncpy(buf, s.c_str(), cap - 1);
        buf[cap - 1] = 0;
    });
}

int rf_query_run(rf_query* q, const rf_arg* args, int64_t n_args, void* output_addr,
                 rf_table** out_table) {
    return guard([&] {
        need(q != nullptr, "query is NULL");
        rf::run_query(q, args, n_args, output_addr, out_table);
    });
}

int rf_query_run_noalloc(rf_query* q, const rf_arg* args, int64_t n_args, void* output_addr,
                         void* const* out_cols, int64_t out_cap, int64_t* out_rows) {
    return guard([&] {
        need(q != nullptr, "query is NULL");
        rf::run_query_noalloc(q, args, n_args, output_addr, out_cols, out_cap, out_rows);
    });
}

int rf_query_time(rf_query* q, const rf_arg* args, int64_t n_args, int32_t iters,
                  double* ms_per_run, double* ms_scan_kernel) {
    return guard([&] {
        need(q && ms_per_run, "NULL argument");
        double scan = 0;
        rf::time_query(q, args, n_args, iters, ms_per_run, &scan);
        if (ms_scan_kernel) *ms_scan_kernel = scan;
    });
}

int rf_batch_create(int32_t max_units, int64_t arena_bytes, int32_t use_graph, rf_batch** out) {
    return guard([&] {
        need(out != nullptr, "NULL argument");
        *out = rf::batch_create(max_units, arena_bytes, use_graph != 0);
    });
}

int rf_batch_footprint(rf_query* q, const rf_arg* args, int64_t n_args, int64_t* bytes) {
    return guard([&] {
        need(q && bytes && (args || n_args == 0), "NULL argument");
        *bytes = rf::batch_footprint(q, args, n_args);
    });
}

int rf_batch_add(rf_batch* b, rf_query* q, const rf_arg* args, int64_t n_args, int32_t* unit) {
    int busy = 0;
    int st = guard([&] {
        need(b && q && unit && (args || n_args == 0), "NULL argument");
        if (!rf::batch_add(b, q, args, n_args, unit)) busy = 1;
    });
    if (st == RF_OK && busy) {
        g_last_error = "round full: no unit slot or arena memory left";
        return RF_ERR_BUSY;
    }
    return st;
}

int rf_batch_launch(rf_batch* b) {
    return guard([&] {
        need(b != nullptr, "NULL argument");
        rf::batch_launch(b);
    });
}

int rf_batch_poll(rf_batch* b, int32_t* done) {
    return guard([&] {
        need(b && done, "NULL argument");
        *done = rf::batch_poll(b) ? 1 : 0;
    });
}

int rf_batch_wait(rf_batch* b) {
    return guard([&] {
        need(b != nullptr, "NULL argument");
        rf::batch_wait(b);
    });
}

int rf_batch_result(rf_batch* b, int32_t unit, void* output_addr, rf_table** out_table) {
    return guard([&] {
        need(b != nullptr, "NULL argument");
        rf::batch_result(b, unit, output_addr, out_table);
    });
}

int rf_batch_reset(rf_batch* b) {
    return guard([&] {
        need(b != nullptr, "NULL argument");
        rf::batch_reset(b);
    });
}

int rf_batch_free(rf_batch* b) {
    return guard([&] { delete b; });
}

int rf_serve(rf_request* reqs, int64_t n, int32_t max_units, int64_t arena_bytes, int32_t use_graph,
             int32_t runners, double* wall_us) {
    return guard([&] {
        need(reqs || n == 0, "NULL argument");
        const double w = rf::serve(reqs, n, max_units, arena_bytes, use_graph != 0, runners);
        if (wall_us) *wall_us = w;
    });
}

int rf_serve_plain(rf_request* reqs, int64_t n, double* wall_us) {
    return guard([&] {
        need(reqs || n == 0, "NULL argument");
        const double w = rf::serve_plain(reqs, n);
        if (wall_us) *wall_us = w;
    });
}

int rf_calib_copy(int64_t bytes, int32_t iters, double* ms_per_copy) {
    return guard([&] {
        need(ms_per_copy && bytes >= 16 && iters > 0, "bad argument");
        rf::Device& d = rf::device();
        int64_t n16 = bytes / 16;
        void* src = rf::dev_alloc((size_t)n16 * 16);
        void* dst = nullptr;
        try {
            dst = rf::dev_alloc((size_t)n16 * 16);
        } catch (...) {
            rf::dev_free(src);
            throw;
        }
        hipEvent_t e0, e1;
        RF_HIP(hipEventCreate(&e0));
        RF_HIP(hipEventCreate(&e1));
        RF_HIP(rf::aot_copy(src, dst, n16, d.stream));
        RF_HIP(hipEventRecord(e0, d.stream));
        for (int i = 0; i < iters; i++) RF_HIP(rf::aot_copy(src, dst, n16, d.stream));
        RF_HIP(hipEventRecord(e1, d.stream));
        RF_HIP(hipEventSynchronize(e1));
        float ms = 0;
        RF_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        rf::dev_free(src);
        rf::dev_free(dst);
        *ms_per_copy = ms / iters;
    });
}

int rf_query_free(rf_query* q) {
    return guard([&] { delete q; });
}

}  // extern "C"
