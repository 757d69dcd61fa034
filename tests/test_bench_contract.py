"""bench.py's command-line contract that holds without a GPU: one GPU only
(any other --gpus value exits with status 2 and says why), and the widened CPU
oracle leg runs on the host alone."""

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_widened_cpu_leg():
    sys.path.insert(0, ROOT)
    import bench

    r = bench.cpu_baseline_widened()
    for key in ("pint_256", "pie_256"):
        assert r[key]["value"] > 0 and r[key]["cores"] == 1 and r[key]["kind"] == "port"
