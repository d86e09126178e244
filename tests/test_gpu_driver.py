"""End-to-end runs of the query driver (rtscan_amd/rts_bindex_driver, the reference
bindex_cuda main over the engine): the reference's own query files on seeded data and a data
file written with rts_data_file_write, one query per launch and batched; every query's
bitmap must equal raw_scan's ([CHECK]hit t/t, no [ERROR]), and the counts must equal numpy's."""
import os
import re
import subprocess

import numpy as np
import pytest

import rtscan_amd as R

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "rtscan_amd", "rts_bindex_driver")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def run(args):
    assert os.path.isfile(DRIVER), "build the driver first (make -C rtscan_amd/csrc)"
    p = subprocess.run([DRIVER] + args, capture_output=True, text=True, timeout=180)
    return p.returncode, p.stdout, p.stderr


def hits(out):
    return [(int(a), int(b)) for a, b in re.findall(r"\[CHECK\]hit (\d+)/(\d+)", out)]


@pytest.mark.parametrize("qfile,cols", [("scan_cmd_32_3c.txt", 3), ("scan_cmd_32_1c.txt", 1),
                                        ("scan_cmd_1e8-2c.txt", 2)])
@pytest.mark.parametrize("mode", ["single", "batch"])
def test_reference_query_files(eng, qfile, cols, mode):
    rc, out, err = run(["-b", str(cols), "-p", os.path.join(GOLDEN, qfile), "-n", "1000003", "-s", "7",
                        "-m", mode])
    assert rc == 0, (rc, err[-2000:], out[-2000:])
    h = hits(out)
    assert h and all(a == b for a, b in h)
    assert "[ERROR]" not in out
    m = re.search(r"\[CHECK\] (\d+) of (\d+) queries matched", out)
    assert m and m.group(1) == m.group(2) and int(m.group(2)) == len(h)


def test_one_column_query_file_in_one_ranges_call(eng):
    """-m ranges: a one-column query file's predicates in one rts_scan_ranges call (the
    reference driver scans them one by one); every bitmap equals raw_scan's."""
    rc, out, err = run(["-b", "1", "-p", os.path.join(GOLDEN, "scan_cmd_32_1c.txt"), "-n", "1000003", "-s", "7",
                        "-m", "ranges"])
    assert rc == 0, (rc, err[-2000:], out[-2000:])
    h = hits(out)
    assert len(h) >= 10 and all(a == b for a, b in h)
    assert "[Time] rts_scan_ranges of %d predicates" % len(h) in out
    assert "[CHECK] %d of %d queries matched" % (len(h), len(h)) in out
    rc, _, err = run(["-b", "3", "-p", os.path.join(GOLDEN, "scan_cmd_32_3c.txt"), "-n", "1000003", "-m", "ranges"])
    assert rc == 2 and "one-column" in err


def test_data_file_and_counts(eng, tmp_path):
    n, cols = 300_007, 2
    data = [R.generate_column(R.DIST_UNIFORM, 3, 0, n),
            R.generate_column(R.DIST_ZIPF, 3, 1, n, 1.5, 2.0 ** 12)]
    path = str(tmp_path / "data.bin")
    arr = (np.concatenate(data)).astype("<u4")
    arr.tofile(path)  # the data-file layout: columns one after another, little-endian
    rng = np.random.default_rng(5)
    lines, want = [], []
    for _ in range(12):
        m = np.ones(n, bool)
        for c in range(cols):
            a, b = sorted(int(v) for v in rng.choice(data[c], 2))
            op = rng.choice(["lt", "le", "gt", "ge", "bt"])
            lines.append("%s %d" % (op, a) + (" %d" % b if op == "bt" else ""))
            x = data[c].astype(np.int64)
            m &= {"lt": x < a, "le": x <= a, "gt": x > a, "ge": x >= a, "bt": (x > a) & (x < b)}[op]
        want.append(int(m.sum()))
    lines.append("exit")
    q = tmp_path / "q.txt"
    q.write_text("\n".join(lines) + "\n")
    for mode in ("single", "batch"):
        rc, out, err = run(["-f", path, "-b", str(cols), "-p", str(q), "-m", mode])
        assert rc == 0, (rc, err[-2000:])
        assert "N = %d" % n in out
        h = hits(out)
        assert [b for _, b in h] == want and all(a == b for a, b in h)


def test_encode_mode_preserves_lt_answers(eng):
    """-e rank-encodes the columns with normalEncode and maps targets with encodeQuery, as
    the reference's ENCODE build does (rtscan.cpp:1016-1026, 1068-1074): every query still
    matches raw_scan on the codes, and because rank codes keep the value order, each `lt`
    query selects as many rows as on the values."""
    q = os.path.join(GOLDEN, "scan_cmd_32_3c.txt")
    args = ["-b", "3", "-p", q, "-n", "500009", "-s", "11"]
    rc0, out0, err0 = run(args)
    rc1, out1, err1 = run(args + ["-e"])
    assert rc0 == 0 and rc1 == 0, (err0[-1000:], err1[-1000:])
    assert "[ENCODE]" in out1 and "distinct values" in out1
    h0, h1 = hits(out0), hits(out1)
    assert all(a == b for a, b in h1)
    assert [b for _, b in h1] == [b for _, b in h0]


def test_encode_mode_targets(eng, tmp_path):
    """encodeQuery's operator rules on real data: le and ge keep the value answers; gt
    takes the start of the next value, so on the codes it misses that value's first row
    (the reference's rule, remap.cpp:256-261); eq / bt cannot be encoded (exit 2)."""
    n = 200_003
    x = R.generate_column(R.DIST_ZIPF, 9, 0, n, 1.5, 2.0 ** 10)
    path = str(tmp_path / "d.bin")
    x.astype("<u4").tofile(path)
    vals = np.unique(x)
    qs = [int(vals[3]), int(vals[len(vals) // 2]), int(vals[-2])]
    lines = []
    for v in qs:
        lines += ["le %d" % v, "ge %d" % v, "gt %d" % v]
    (tmp_path / "q.txt").write_text("\n".join(lines + ["exit"]) + "\n")
    rc, out, err = run(["-f", path, "-b", "1", "-p", str(tmp_path / "q.txt"), "-e"])
    assert rc == 0, err[-1000:]
    h = hits(out)
    want = []
    for v in qs:
        want += [int((x <= v).sum()), int((x >= v).sum()), int((x > v).sum()) - 1]
    assert [b for _, b in h] == want and all(a == b for a, b in h)
    (tmp_path / "e.txt").write_text("eq 5\nexit\n")
    rc, out, err = run(["-f", path, "-b", "1", "-p", str(tmp_path / "e.txt"), "-e"])
    assert rc == 2 and "cannot be encoded" in err


def test_target_lists_follow_the_reference_checker(eng, tmp_path):
    """Query lines with comma-separated targets (the grammar of bindex.cpp:2168-2189): each
    column selects the rows hit by an odd number of its targets, as the reference's checker
    XORs every target's hits (raw_scan_entry, bindex.cpp:1894-1930); mixed with ordinary
    queries, one query per launch and batched, against numpy."""
    n, cols = 200_003, 2
    data = [R.generate_column(R.DIST_UNIFORM, 4, 0, n), R.generate_column(R.DIST_ZIPF, 4, 1, n, 1.1, 2.0 ** 10)]
    path = str(tmp_path / "data.bin")
    np.concatenate(data).astype("<u4").tofile(path)
    rng = np.random.default_rng(9)
    ops = {"lt": lambda x, a, b: x < a, "le": lambda x, a, b: x <= a, "gt": lambda x, a, b: x > a,
           "ge": lambda x, a, b: x >= a, "eq": lambda x, a, b: x == a,
           "bt": lambda x, a, b: (x > min(a, b)) & (x < max(a, b))}
    lines, want = [], []
    for qi in range(10):
        m = np.ones(n, bool)
        for c in range(cols):
            op = str(rng.choice(list(ops)))
            k = 1 if qi % 3 == 0 else int(rng.integers(2, 4))
            t1 = [int(v) for v in rng.choice(data[c], k)]
            t2 = [int(v) for v in rng.choice(data[c], k)]
            x = data[c].astype(np.int64)
            col = np.zeros(n, bool)
            for a, b in zip(t1, t2):
                col ^= ops[op](x, a, b)
            m &= col
            lines.append("%s %s" % (op, ",".join(map(str, t1))) + (" " + ",".join(map(str, t2)) if op == "bt" else ""))
        want.append(int(m.sum()))
    lines.append("exit")
    q = tmp_path / "q.txt"
    q.write_text("\n".join(lines) + "\n")
    for mode in ("single", "batch"):
        rc, out, err = run(["-f", path, "-b", str(cols), "-p", str(q), "-m", mode])
        assert rc == 0, (rc, err[-2000:], out[-3000:])
        h = hits(out)
        assert [b for _, b in h] == want and all(a == b for a, b in h), (mode, h, want)
