"""The benchmark's command-line contract: one GPU only (--gpus N > 1 exits 2 with a message), the
HBM traffic figure it reads from the committed counter profiles, and (GPU) one JSON line with
the fields the benchmark contract names, for a short run of the headline workload."""
import importlib
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args, timeout=300):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *map(str, args)], capture_output=True,
                          text=True, timeout=timeout, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", 2, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_traffic_read_from_the_committed_profiles():
    bench = importlib.import_module("bench")
    n, c, h, k = 32, 256, 56, 256
    algorithmic = 2 * n * c * h * h + 32 * k * c + 2 * n * k * h * h  # input + U + output bytes
    fused = bench.measured_traffic("fused", "form=fused")
    assert fused is not None and algorithmic <= fused <= 2 * algorithmic
    # a key for another kernel variant than the profile's gives no figure
    assert bench.measured_traffic("fused", "cfg=9") is None
    staged = bench.measured_traffic("staged", "cfg=9")
    assert staged is not None and staged > 32 * n * c * (h // 2) ** 2  # at least V's bytes


@pytest.mark.gpu
def test_bench_line_contract(cuda):
    print("case bench.py --steps 5 --warmup 1 --no-forms --no-cpu", flush=True)
    r = _bench("--steps", 5, "--warmup", 1, "--no-forms", "--no-cpu")
    assert r.returncode == 0, r.stderr[-2000:]
    lines = r.stdout.strip().splitlines()
    assert len(lines) == 1
    line = json.loads(lines[0])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
                "scaling", "vs_baseline", "dtype", "data", "config", "roofline"):
        assert key in line, key
    assert line["n_gpus"] == 1 and line["steps"] == 5 and line["warmup"] == 1
    assert line["value"] > 0 and line["higher_is_better"] is True and line["vs_baseline"] is None
    assert line["dtype"] == "fp16" and "workload" in line["config"]
    roof = line["roofline"]
    for key in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert key in roof, key
    assert 0 < roof["frac"] < 1 and abs(roof["frac"] - roof["achieved"] / roof["peak"]) < 1e-3
    # direct-conv FLOPs of the workload over the step time give the value
    flops = 2.0 * 32 * 256 * 256 * 9 * 56 * 56
    assert abs(line["value"] - flops / (line["ms_per_step"] * 1e-3) / 1e12) <= 0.01 * line["value"]
    err = line["error_vs_fp64"]
    assert err["finite"] and err["max_rel"] <= 1e-2
    print(f"  value {line['value']} TFLOP/s, frac {roof['frac']}", flush=True)
