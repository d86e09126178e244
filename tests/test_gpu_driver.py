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
