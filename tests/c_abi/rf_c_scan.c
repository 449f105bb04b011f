/* A C host of the library's C ABI, with no Python: builds a record batch
 * through the Arrow C data interface, uploads it, compiles a query from its
 * IR text, runs it and checks the result against a plain C loop. The
 * query is the C1 scan-filter-aggregate (COUNT, SUM(v*k) WHERE lo <= d < hi
 * AND v < vmax over T(rowid i64, k i64, v f64, d i32), dim 2), the shape of
 * the reference's TPC-H Q6 call (benchmark/include/tpch.h:161-208).
 *
 * usage: rf_c_scan IR_FILE LO HI VMAX [ROWS]   (exit 0 on a match) */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "reffine_hip.h"

static uint64_t lcg(uint64_t* s) {
    *s = *s * 6364136223846793005ull + 1442695040888963407ull;
    return *s >> 11;
}

static int check(int st, const char* what) {
    if (st != RF_OK) {
        fprintf(stderr, "%s failed: status %d: %s\n", what, st, rf_last_error());
        exit(1);
    }
    return st;
}

int main(int argc, char** argv) {
    if (argc < 5) {
        fprintf(stderr, "usage: %s IR_FILE LO HI VMAX [ROWS]\n", argv[0]);
        return 2;
    }
    FILE* f = fopen(argv[1], "rb");
    if (!f) { perror(argv[1]); return 2; }
    fseek(f, 0, SEEK_END);
    long len = ftell(f);
    fseek(f, 0, SEEK_SET);
    char* ir = malloc((size_t)len + 1);
    if (fread(ir, 1, (size_t)len, f) != (size_t)len) { fclose(f); return 2; }
    ir[len] = 0;
    fclose(f);
    const int32_t lo = atoi(argv[2]), hi = atoi(argv[3]);
    const double vmax = atof(argv[4]);
    const int64_t n = argc > 5 ? atoll(argv[5]) : (1 << 20);

    /* Columns. */
    int64_t* rowid = malloc((size_t)n * 8);
    int64_t* k = malloc((size_t)n * 8);
    double* v = malloc((size_t)n * 8);
    int32_t* d = malloc((size_t)n * 4);
    uint64_t s = 42;
    for (int64_t i = 0; i < n; i++) {
        rowid[i] = i;
        k[i] = (int64_t)(lcg(&s) % 1000);
        v[i] = (double)(lcg(&s) >> 0) * (1.0 / 9007199254740992.0);
        d[i] = (int32_t)(lcg(&s) % 2556);
    }

    /* Arrow C data interface: a struct array with one child per column. */
    const char* formats[4] = {"l", "l", "g", "i"};
    const void* values[4] = {rowid, k, v, d};
    struct ArrowSchema cs[4], root_s;
    struct ArrowSchema* cs_ptr[4];
    struct ArrowArray ca[4], root_a;
    struct ArrowArray* ca_ptr[4];
    const void* bufs[4][2];
    const void* root_bufs[1] = {NULL};
    for (int c = 0; c < 4; c++) {
        memset(&cs[c], 0, sizeof cs[c]);
        cs[c].format = formats[c];
        cs[c].name = "";
        cs_ptr[c] = &cs[c];
        memset(&ca[c], 0, sizeof ca[c]);
        bufs[c][0] = NULL;
        bufs[c][1] = values[c];
        ca[c].length = n;
        ca[c].n_buffers = 2;
        ca[c].buffers = bufs[c];
        ca_ptr[c] = &ca[c];
    }
    memset(&root_s, 0, sizeof root_s);
    root_s.format = "+s";
    root_s.name = "";
    root_s.n_children = 4;
    root_s.children = cs_ptr;
    memset(&root_a, 0, sizeof root_a);
    root_a.length = n;
    root_a.n_buffers = 1;
    root_a.buffers = root_bufs;
    root_a.n_children = 4;
    root_a.children = ca_ptr;

    check(rf_init(0), "rf_init");
    rf_table* t = NULL;
    check(rf_table_upload(&root_s, &root_a, 2, &t), "rf_table_upload");
    rf_query* q = NULL;
    check(rf_compile(ir, 0, &q), "rf_compile");

    /* Errors come back as status codes with a message. */
    rf_query* bad = NULL;
    if (rf_compile("(func", 0, &bad) == RF_OK || strlen(rf_last_error()) == 0) {
        fprintf(stderr, "malformed IR was accepted\n");
        return 1;
    }

    struct { int64_t cnt; double sum; } out = {0, 0.0};
    rf_arg a;
    memset(&a, 0, sizeof a);
    a.kind = 0;
    a.u.table = t;
    check(rf_query_run(q, &a, 1, &out, NULL), "rf_query_run");

    int64_t cnt = 0;
    double sum = 0.0;
    for (int64_t i = 0; i < n; i++)
        if (d[i] >= lo && d[i] < hi && v[i] < vmax) {
            cnt++;
            sum += v[i] * (double)k[i];
        }
    const double rel = fabs(out.sum - sum) / (fabs(sum) > 0 ? fabs(sum) : 1.0);
    printf("c abi: rows %lld, count %lld (C loop %lld), sum %.17g (C loop %.17g, rel %.2e)\n",
           (long long)n, (long long)out.cnt, (long long)cnt, out.sum, sum, rel);
    check(rf_query_free(q), "rf_query_free");
    check(rf_table_free(t), "rf_table_free");
    free(ir); free(rowid); free(k); free(v); free(d);
    return (out.cnt == cnt && rel <= 1e-12) ? 0 : 1;
}
