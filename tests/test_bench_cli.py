"""CPU checks of bench.py's command-line contract (no GPU work is started)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_exits_with_a_message():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_unknown_extras_leg_is_rejected():
    r = _bench("--extras", "sizes,nonsense", "--no-cpu")
    assert r.returncode == 2
    assert "nonsense" in r.stderr
