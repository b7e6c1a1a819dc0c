"""bench.py's command line on the CPU: the one-GPU contract refuses more GPUs
(and a multi-rank launch) with a message and exit status 2, before anything
touches a GPU."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120, env=e)


def test_more_than_one_gpu_is_refused():
    p = _bench(["--gpus", "2"])
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout == ""


def test_multi_rank_launch_is_refused():
    p = _bench(["--gpus", "1"], {"WORLD_SIZE": "4"})
    assert p.returncode == 2
    assert "one GPU" in p.stderr
