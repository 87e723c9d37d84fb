"""bench.py's command line without a GPU: it measures one GPU, and a request for more exits with
status 2 and a message saying so, before any device work."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    done = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1", "--warmup", "0"],
                          capture_output=True, text=True, timeout=120)
    assert done.returncode == 2, done.stdout + done.stderr
    assert "measures one GPU" in done.stdout and "--gpus 2" in done.stdout, done.stdout
