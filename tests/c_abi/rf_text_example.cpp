// A reffine program on the GPU path: the scan-filter-aggregate
// COUNT(*), SUM(v * k) WHERE 365 <= d < 730 AND v < 0.5 over T(rowid, k, v, d)
// built with reffine's own builder (include/reffine/builder/reffiner.h),
// written as IR text by reffine_hip::to_rf_text and compiled by rf_compile.
// Compiled -fsyntax-only against the reference's headers by the tests (it
// links against reffine's IR library and libreffine_hip in a reffine build).
#include <cstdio>

#include "reffine/builder/reffiner.h"
#include "reffine_hip_text.hpp"

using namespace reffine;
using namespace reffine::reffiner;

std::shared_ptr<Func> scan_c1()
{
    auto T = _sym("T", DataType(BaseType::VECTOR, {_i64_t, _i64_t, _f64_t, _i32_t}, 2,
                                {EncodeType::FLAT, EncodeType::FLAT, EncodeType::FLAT,
                                 EncodeType::FLAT}));
    auto red = _red(
        _subvec(T, _idx(0), _len(T, 1)),
        []() { return _new(vector<Expr>{_i64(0), _f64(0)}); },
        [](Expr s, Expr v) {
            auto k = _get(v, 0), x = _get(v, 1), d = _get(v, 2);
            auto pred = _gte(d, _i32(365)) & _lt(d, _i32(730)) & _lt(x, _f64(0.5));
            auto next = _new(vector<Expr>{_add(_get(s, 0), _i64(1)),
                                          _add(_get(s, 1), _mul(x, _cast(_f64_t, k)))});
            return _sel(pred, next, s);
        });
    auto red_sym = _sym("red", red);
    auto fn = _func("scan_c1", red_sym, vector<Sym>{T});
    fn->tbl[red_sym] = red;
    return fn;
}

struct C1Out {
    int64_t count;
    double sum;
};

int run_scan_c1(::ArrowTable* t, C1Out* out)
{
    if (rf_init(0) != RF_OK) return 1;
    std::printf("%s\n", reffine_hip::to_rf_text(scan_c1()).c_str());
    reffine_hip::HipQuery q(scan_c1());        // compile_op(...)
    reffine_hip::Table tbl(t);                 // the ArrowTable in HBM
    q(out, {reffine_hip::table(tbl)});         // fn(&out, t)
    return 0;
}
