"""bench.py's command line without a GPU: it runs on one GPU only and says so
(exit status 2, no JSON line) when asked for more."""

import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("n", [2, 8])
def test_more_than_one_gpu_is_refused(n):
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", str(n)],
                       capture_output=True, text=True, timeout=300, cwd=REPO)
    assert r.returncode == 2
    assert "exactly one GPU" in r.stdout + r.stderr
    assert '{"metric"' not in r.stdout
