"""Host code under the sanitizers (CPU only): the request loop's admission logic
(rtscan_amd/csrc/admission.h) built with AddressSanitizer + UndefinedBehaviorSanitizer and
with ThreadSanitizer, and the oracle's C restatement built with ASan + UBSan.  Each program
is compiled with g++/gcc on the host (no device code) and must exit 0 with no report."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.join(ROOT, "tests", "host_logic")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


def build_and_run(tmp_path, compiler, sources, flags, name, env_extra=None):
    exe = str(tmp_path / name)
    cmd = [compiler, "-O1", "-g", "-fno-omit-frame-pointer", "-o", exe] + sources + flags
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = dict(os.environ)
    env.update(env_extra or {})
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "ERROR: " not in r.stderr and "WARNING: ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    return r.stdout


def test_admission_asan_ubsan(tmp_path):
    out = build_and_run(tmp_path, "g++", [os.path.join(HERE, "admission_test.cpp")],
                        ["-std=c++17", "-pthread", "-I" + os.path.join(ROOT, "rtscan_amd", "csrc"),
                         "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"],
                        "admission_asan", {"ASAN_OPTIONS": "detect_leaks=1"})
    assert "admission ok" in out


def test_admission_tsan(tmp_path):
    out = build_and_run(tmp_path, "g++", [os.path.join(HERE, "admission_test.cpp")],
                        ["-std=c++17", "-pthread", "-I" + os.path.join(ROOT, "rtscan_amd", "csrc"),
                         "-fsanitize=thread"],
                        "admission_tsan", {"TSAN_OPTIONS": "halt_on_error=1"})
    assert "admission ok" in out


def test_oracle_asan_ubsan(tmp_path):
    out = build_and_run(tmp_path, "gcc", [os.path.join(HERE, "oracle_sanitize.c"),
                                          os.path.join(ROOT, "oracle", "oracle.c")],
                        ["-std=c11", "-fopenmp", "-I" + os.path.join(ROOT, "oracle"),
                         "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"],
                        "oracle_asan", {"ASAN_OPTIONS": "detect_leaks=1"})
    assert "oracle ok" in out
