"""The benchmark's default workload as the round-end run invokes it, shortened: it must
print one JSON line with the metric, the roofline and the learned report, and exit 0."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.timeout(600)
def test_bench_default_workload_prints_a_full_line():
    print("case bench.py --warmup 5 --steps 2 --no-plain --no-cpu-baseline --no-accuracy", flush=True)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--warmup", "5", "--steps", "2",
                        "--no-plain", "--no-cpu-baseline", "--no-accuracy"],
                       capture_output=True, text=True, timeout=540, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-3000:]
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    assert len(lines) == 1, r.stdout[-2000:]
    out = json.loads(lines[0])
    assert out["metric"].startswith("body-substeps/s on 1 MI355X")
    assert out["value"] > 0 and out["steps"] == 2 and out["warmup"] == 5 and out["dtype"] == "f32"
    assert out["config"]["bodies"] == 262144 and out["config"]["scenes"] == 1024
    assert out["roofline"] is not None and out["roofline"]["frac"] > 0
    assert out["learned"] is not None
    assert out["learned"].get("logged_samples", 0) > 0 or out["learned"]["value"] is None
