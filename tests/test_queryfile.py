"""The reference's query-file and data-file formats, parsed by the library's host code
(rts_query.h), checked on the reference's own query files (tests/golden/, copied from the
reference's test/ directory) and on crafted edge cases."""
import os

import numpy as np
import pytest

import rtscan_amd as R

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_scan_cmd_32_1c():
    # block 1 of test/scan_cmd_32_1c.txt: 11 lt queries at round(i * (2^32 - 1) / 10), the
    # first written as lt 1; everything after the first exit is ignored
    qs = R.parse_query_file(path=os.path.join(GOLDEN, "scan_cmd_32_1c.txt"), ncols=1)
    assert len(qs) == 11
    want = [1] + [int(round(i * (2 ** 32 - 1) / 10)) for i in range(1, 11)]
    got = [q[0][1][0] for q in qs]
    assert all(q[0][0] == R.LT and q[0][2] is None for q in qs)
    assert [abs(a - b) <= 1 for a, b in zip(got, want)] == [True] * 11
    assert got[-1] == 4294967295


def test_multi_column_files():
    qs = R.parse_query_file(path=os.path.join(GOLDEN, "scan_cmd_32_3c.txt"), ncols=3)
    assert len(qs) == 11 and all(len(q) == 3 for q in qs)
    qs = R.parse_query_file(path=os.path.join(GOLDEN, "scan_cmd_1e8-2c.txt"), ncols=2)
    assert [q[0][1][0] for q in qs] == [i * 10000000 for i in range(11)]
    qs = R.parse_query_file(path=os.path.join(GOLDEN, "zipf1.5.txt"), ncols=3)
    assert qs == [[(R.LT, [59], None)] * 3]
    qs = R.parse_query_file(path=os.path.join(GOLDEN, "normal.txt"), ncols=3)
    assert qs == [[(R.LT, [2147483650], None)] * 3]


def test_reference_file_with_a_comment_line_is_rejected():
    """test/uniform0.9.txt has `# uniform0.9` as its 4th line, before any exit.  The
    reference parses `uniform0.9` with stod (bindex.cpp:1845), which throws and ends the
    process; the library returns a syntax error naming the line instead."""
    with pytest.raises(R.QueryFileError, match="line 4"):
        R.parse_query_file(path=os.path.join(GOLDEN, "uniform0.9.txt"), ncols=3)


def test_grammar_edge_cases():
    qs = R.parse_query_file("bt 10 5\neq 7\nexit\ngarbage after exit\n", ncols=2)
    assert qs == [[(R.BT, [10], [5]), (R.EQ, [7], None)]]
    qs = R.parse_query_file("lt 1,2,3\nexit\n", ncols=1)
    assert qs == [[(R.LT, [1, 2, 3], None)]]
    qs = R.parse_query_file("bt 1,5 9,20\r\nexit\r\n", ncols=1)
    assert qs == [[(R.BT, [1, 5], [9, 20])]]
    qs = R.parse_query_file("le 1.9e3\nexit", ncols=1)  # stod then cast (bindex.cpp:1833)
    assert qs == [[(R.LE, [1900], None)]]
    assert R.parse_query_file("exit\n", ncols=3) == []
    assert R.parse_query_file("", ncols=1) == []


@pytest.mark.parametrize("text", ["xx 5\nexit", "lt\nexit", "lt -1\nexit", "lt 4294967296\nexit",
                                  "bt 5\nexit", "lt 5 6\nexit", "lt 1,,2\nexit", "lt 5\nexit",
                                  "lt 5  6\nexit"])
def test_grammar_errors(text):
    with pytest.raises(R.QueryFileError):
        R.parse_query_file(text, ncols=2 if text == "lt 5\nexit" else 1)


def test_missing_file():
    with pytest.raises(R.QueryFileError):
        R.parse_query_file(path="/nonexistent/q.txt", ncols=1)


def test_data_file_round_trip(tmp_path):
    cols = [R.generate_column(R.DIST_UNIFORM, 1, c, 1000) for c in range(3)]
    p = str(tmp_path / "data.dat")
    R.write_data_file(p, cols)
    raw = np.fromfile(p, dtype="<u4")
    assert np.array_equal(raw, np.concatenate(cols))  # column after column
    back = R.read_data_file(p, 3, 1000)
    assert all(np.array_equal(a, b) for a, b in zip(cols, back))
    with pytest.raises(OSError):
        R.read_data_file(p, 4, 1000)
