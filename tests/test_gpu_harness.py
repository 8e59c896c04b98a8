"""GPU tests of the C++ harness lib/conv2dfp16demo, the counterpart of the reference harness
(src/main.cpp:8-166): a separately compiled caller that calls getParamsize / getkernelInfo and
launches the four slots itself with hipExtLaunchKernel (main.cpp:105-108), then checks the
output against the FP64 direct convolution and the reference's own elementwise criterion
(main.cpp:140-151)."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu


def _run(wc, args):
    assert os.path.exists(wc.DEMO_PATH), "harness not built"
    print(f"case harness {' '.join(map(str, args))}", flush=True)
    r = subprocess.run([wc.DEMO_PATH, *map(str, args)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    line = json.loads(r.stdout.strip().splitlines()[-1])
    print(f"  {line}", flush=True)
    return line


def test_harness_cfg1_positive_meets_reference_criterion(wc, cuda):
    """N=1 C=64 56x56 K=64 on the reference's positive input range: no element fails the
    reference's |gpu - cpu| / gpu <= 1% check, and the normwise error is within 1e-2."""
    line = _run(wc, [1, 64, 56, 56, 64, 3, 3, 1, 1, 1, 1, "--dist", "positive", "--iters", "5"])
    assert line["method"] == "winograd" and line["form"] == "fused"
    assert line["reference_criterion_errors"] == 0
    assert line["max_rel"] <= 1e-2


def test_harness_ragged_cfg2_variant(wc, cuda):
    """N=32 C=K=96 55x55 (partial tiles, channels padded to 32) through the caller-side slots."""
    line = _run(wc, [32, 96, 55, 55, 96, 3, 3, 1, 1, 1, 1, "--iters", "5"])
    assert line["method"] == "winograd"
    assert line["max_rel"] <= 1e-2 and line["rms_rel"] <= 1e-2


def test_harness_staged_and_direct_slots(wc, cuda):
    """A staged-form layer (K = 512 with a large N*T*K*C) and a general-geometry layer (5x5,
    stride 2, pad 2: the direct path's four slots) through the same caller-side launches."""
    staged = _run(wc, [8, 256, 56, 56, 512, 3, 3, 1, 1, 1, 1, "--iters", "3"])
    assert staged["method"] == "winograd" and staged["form"] == "staged"
    assert staged["max_rel"] <= 1e-2
    direct = _run(wc, [2, 16, 20, 20, 24, 5, 5, 2, 2, 2, 2, "--dist", "positive", "--iters", "3"])
    assert direct["method"] == "direct"
    assert direct["max_rel"] <= 1e-2 and direct["reference_criterion_errors"] == 0
