"""Command-line contract of bench.py that holds without a GPU.

The bench measures one GPU: asked for more it must stop at once with a
message and a non-zero status, before it touches a device, so that a
launcher never records a line for a configuration that was not run.
"""

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          cwd=ROOT, capture_output=True, text=True, timeout=300)


def test_more_than_one_gpu_is_refused_with_a_message():
    for n in ("2", "8"):
        p = _bench("--gpus", n, "--steps", "1", "--warmup", "0")
        assert p.returncode == 2, (p.returncode, p.stderr[-500:])
        msg = p.stdout + p.stderr
        assert "one GPU only" in msg
        # no JSON line: nothing was measured
        assert not any(line.lstrip().startswith("{") for line in p.stdout.splitlines())
