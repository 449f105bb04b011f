"""The reference test suite's queries (test/src/test_reduce.cpp,
test_transform.cpp, test_join.cpp, test_utils.cpp) through the GPU path,
against the values those tests assert (tests/golden/expected.json)."""
import numpy as np
import pytest

import ref_queries as rq
from reffine_amd import DeviceTable, compile_op

pytestmark = pytest.mark.gpu


def students_table(dim=1):
    return DeviceTable.from_columns(rq.students(), dim=dim, names=rq.STUDENT_COLS)


@pytest.mark.parametrize("vectorize", [False, True])
@pytest.mark.parametrize("det", [False, True])
def test_aggregate_golden(device, vectorize, det):
    exp = rq.expected()
    t2 = students_table(2)
    assert compile_op(rq.sum_hours_subvec(), vectorize, det)(t2) == exp["aggregate_loop_test"]
    t1 = students_table(1)
    assert compile_op(rq.vector_op(), vectorize, det)(t1) == exp["aggregate_op_test"]


@pytest.mark.parametrize("vectorize", [False, True])
def test_transform_golden(device, vectorize):
    exp = rq.expected()
    ids, hours = rq.students()[0], rq.students()[1]
    out = compile_op(rq.transform_op(), vectorize)(students_table())
    assert out.nrows == exp["transform_op_test_len"]
    np.testing.assert_array_equal(out.column(0), ids[5:])
    np.testing.assert_array_equal(out.column(1), hours[5:] + 10)


def test_gen_table_and_join_golden(device):
    exp = rq.expected()
    gen = compile_op(rq.gen_table_op())
    tabs = []
    for lb, ub in [(0, 10), (5, 15), (7, 12)]:
        t = gen(lb, ub)
        assert t.nrows == ub - lb
        np.testing.assert_array_equal(t.column(0), np.arange(lb, ub))
        np.testing.assert_array_equal(t.column(1), np.arange(lb, ub))
        tabs.append(t)
    out = compile_op(rq.join_op())(*tabs)
    np.testing.assert_array_equal(out.column(0), exp["join_op_test_keys"])
    np.testing.assert_array_equal(out.column(1), 3 * np.array(exp["join_op_test_keys"]))


def test_empty_range_and_single_row(device):
    gen = compile_op(rq.gen_table_op())
    assert gen(5, 5).nrows == 0
    assert gen(9, 3).nrows == 0
    t = gen(-3, -2)
    np.testing.assert_array_equal(t.column(0), [-3])


@pytest.mark.parametrize("a_ub,b_ub", [(10, 10), (3, 7), (1, 1), (0, 5)])
def test_nested_golden(device, a_ub, b_ub):
    """test/src/test_nested.cpp:42-53: row idx holds (idx / b_ub, idx % b_ub,
    their sum), a outer and b inner."""
    out = compile_op(rq.nested_op(a_ub, b_ub))()
    n = a_ub * b_ub
    assert out.nrows == n
    idx = np.arange(n)
    assert np.array_equal(out.column(0), idx // max(b_ub, 1))
    assert np.array_equal(out.column(1), idx % max(b_ub, 1))
    assert np.array_equal(out.column(2), idx // max(b_ub, 1) + idx % max(b_ub, 1))
