"""bench.py contract checks that need no GPU: one GPU only, so any --gpus
other than 1 ends with a JSON error line and exit status 2 before the GPU or
the FP64 reference is touched."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("extra", [[], ["--config", "c4"], ["--config", "c5"]])
def test_more_than_one_gpu_is_refused(extra):
    cp = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", *extra],
                        capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert cp.returncode == 2, cp.stderr
    line = cp.stdout.strip().splitlines()[-1]
    assert "one GPU" in json.loads(line)["error"]
