"""bench.py's output contract on a short run: one JSON line on stdout with
the metric, value and unit, the step counts it was asked for, a roofline
object whose achieved/peak/frac agree, a cpu_baseline object, and the
metric's accuracy condition (<= 1e-6 per particle against the reference's
C rows) met by the timed steps."""
import json
import os
import subprocess
import sys

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_bench_json_line():
    cmd = [sys.executable, os.path.join(ROOT, "bench.py"), "--steps", "3",
           "--warmup", "1", "--cpu-seconds", "0.5"]
    print("[bench contract] " + " ".join(cmd), flush=True)
    p = subprocess.run(cmd, cwd=ROOT, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, universal_newlines=True,
                       timeout=240)
    assert p.returncode == 0, p.stderr[-2000:]
    lines = [l for l in p.stdout.splitlines() if l.strip()]
    assert len(lines) == 1, lines
    r = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup",
              "ms_per_step", "higher_is_better", "scaling", "vs_baseline",
              "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in r, k
    assert r["n_gpus"] == 1 and r["steps"] == 3 and r["warmup"] == 1
    assert r["value"] > 0 and r["higher_is_better"] is True
    assert "workload" in r["config"]
    rf = r["roofline"]
    assert rf["bound"] in ("valu", "hbm", "mfma")
    assert abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-9
    assert 0 < rf["frac"] < 1
    cb = r["cpu_baseline"]
    for k in ("value", "unit", "cores", "kind", "sample"):
        assert k in cb, k
    assert cb["kind"] in ("reference", "port") and cb["cores"] >= 1
    assert cb["fp64_ieee_bitwise_equal"] is True
    # the metric's accuracy condition, against the reference's C rows
    acc = r["accuracy"]
    assert acc["rows"] >= 64
    assert acc["acc_max"] <= 1e-6 and acc["jerk_max"] <= 1e-6, acc
    assert r["accuracy_vs_fp64_gpu"]["jerk_max"] <= 1e-6
    assert r["fp32c_rows_recomputed_fp64"] > 0
