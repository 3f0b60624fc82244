"""bench.py's command line: more than one GPU is refused with a message and
exit status 2 before any device work (this build measures one GPU)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args, env=None):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True,
                          text=True, timeout=120, env=env, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    out = _bench("--gpus", "2")
    assert out.returncode == 2
    assert "one GPU" in out.stderr and out.stdout == ""


def test_world_size_above_one_is_refused():
    env = dict(os.environ, WORLD_SIZE="4", RANK="0", LOCAL_RANK="0")
    out = _bench("--gpus", "1", env=env)
    assert out.returncode == 2
    assert "one GPU" in out.stderr
