"""bench.py end to end at a reduced size: one JSON line carrying the keys of
the benchmark contract, the roofline object priced on the FP64 vector
peak, and the CPU leg (reference CPU path, all-cores figure, oracle
precision sample, track-record check).  The batched, encounter and
end-to-end legs are left out to keep the run short; they have their own
GPU tests."""
import json
import subprocess
import sys

import pytest

from tests.conftest import ROOT

pytestmark = pytest.mark.gpu

KEYS = ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
        "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline")


def test_bench_json_line():
    cmd = [sys.executable, str(ROOT / "bench.py"), "--particles", "20000", "--steps", "2", "--warmup", "1",
           "--no-batch", "--no-encounter", "--no-config0", "--cpu-sample", "300", "--cpu-threads", "2",
           "--oracle-sample", "16"]
    print(" ".join(cmd[1:]), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=240, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-3000:]
    lines = [line for line in r.stdout.splitlines() if line.startswith("{")]
    assert len(lines) == 1, r.stdout[-2000:]
    d = json.loads(lines[0])
    assert all(k in d for k in KEYS), [k for k in KEYS if k not in d]
    assert d["n_gpus"] == 1 and d["steps"] == 2 and d["warmup"] == 1 and d["higher_is_better"] is True
    assert d["value"] > 0 and d["ms_per_step"] > 0 and d["dtype"] == "f64+f32"
    assert set(d["forms"]) == {"fp64", "fast", "mixed"}
    assert d["value"] == d["forms"]["mixed"]["value"]
    rf = d["roofline"]
    assert rf["peak"] == 78.6 and rf["unit"] == "TFLOP/s" and 0.0 < rf["frac"] < 1.0
    assert abs(rf["achieved"] / rf["peak"] - rf["frac"]) < 1e-12
    cb = d["cpu_baseline"]
    assert cb["cores"] == 1 and cb["value"] > 0 and cb["kind"] in ("reference", "port")
    assert cb["all_cores"]["cores"] == 2 and cb["all_cores"]["value"] > 0
    po = cb["precision_vs_oracle"]
    assert po["sample"] == 16 and po["after_steps"] == 3 * 2048
    assert all(po[f]["flags_differ"] == 0 and po[f]["da_over_a"]["max"] < 1e-10 for f in ("fp64", "fast", "mixed"))
    tr = d["track"]
    assert tr["particles"] == 20000 and tr["file_bytes"] > 20000 * 28
    if cb["kind"] == "reference":
        # device element records = the reference writer's bytes
        assert cb["track"]["ids_equal"] and cb["track"]["max_ulps"] == 0 and cb["track"]["floats_equal"] == 1.0
