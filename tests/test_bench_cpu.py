"""bench.py's host arithmetic: TTS99, the roofline entry and the one-GPU refusal (no GPU needed)."""
import math
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


@pytest.mark.parametrize("p", [0.01, 0.125, 0.25, 0.5, 0.9, 0.999])
def test_tts99_is_the_read_count_for_99_percent(p):
    """TTS99 / t_read reads leave a 1 % chance of no success: (1 - p) ** k = 0.01
    (examples/python/dwave_comparative_benchmark.py:438-443)."""
    t = 1.7
    k = bench.tts99(t, p) / t
    assert math.isclose((1.0 - p) ** k, 0.01, rel_tol=1e-12)


def test_tts99_edges_and_order():
    assert bench.tts99(0.3, 0.0) is None          # no success: no finite TTS99
    assert bench.tts99(0.3, 1.0) == 0.3           # every read succeeds: one read
    vals = [bench.tts99(1.0, p) for p in (0.05, 0.2, 0.6, 0.95)]
    assert all(a > b for a, b in zip(vals, vals[1:]))  # more success, less time
    # the lower bound bench reports for a method with no success in r reads: one success in r + 1
    r = 8
    assert math.isclose(bench.tts99(1.172, 1.0 / (r + 1)), 1.172 * math.log(0.01) / math.log(r / (r + 1)))


def test_roofline_fraction():
    roof = bench._roof(4e12, 1.0)               # 4 TB in one second
    assert roof["bound"] == "hbm" and roof["unit"] == "GB/s" and roof["peak"] == 8000.0
    assert math.isclose(roof["achieved"], 4000.0) and math.isclose(roof["frac"], 0.5)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""
