"""bench.py's command line without a GPU: more than one GPU is refused with a message
before any device is touched, and the workload options are listed."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def run(args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([sys.executable, BENCH] + args, capture_output=True, text=True,
                          timeout=120, env=e)


def test_more_than_one_gpu_is_refused():
    p = run(["--gpus", "2"])
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""


def test_multi_rank_launch_is_refused():
    p = run([], env={"WORLD_SIZE": "2"})
    assert p.returncode == 2 and "one GPU" in p.stderr


def test_help_lists_the_workloads():
    p = run(["--help"])
    assert p.returncode == 0
    for w in ("scan", "conj", "batch", "serve", "config1", "readprobe", "--batch-config", "--steps",
              "--warmup"):
        assert w in p.stdout
