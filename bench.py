"""Benchmark: plain FP16 Winograd F(2x2,3x3) conv2d on one MI355X.

Workload (one "step" = one forward pass over one batch): N=32, C=256, H=W=56, K=256,
3x3 stride 1 pad 1, FP16 NCHW input ~N(0,1) and KCRS filters ~N(0, 2/(9C)), seeded.
A step runs the reference's four-slot protocol (filter transform, input transform,
MFMA GEMM with fused output transform, empty fourth slot), as the reference harness
times it (main.cpp:112-131). Inputs are resident in HBM before timing starts.

value = effective TFLOP/s = 2*N*K*C*9*Oh*Ow / time per step (direct-conv FLOPs).
Also reported: per-stage times, the steady state (weights prepared once), the
roofline of the dominant kernel (the GEMM), the error against the FP64 direct
convolution (GPU validation kernel), and the reference's CPU checker timed on the
host on one image of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Eff. conv TFLOP/s & % MFMA peak: plain fp16 vs fp8/batched at rel.err≤1e-2"
PEAK_FP16_TFLOPS = 2516.6  # MI355X dense FP16 MFMA: 256 CU x 2.4 GHz x 4096 FLOP/clk/CU
PEAK_HBM_GBS = 8000.0

WORKLOAD = dict(n=32, c=256, h=56, w=56, k=256)


def _events_time(fn, iters: int) -> float:
    """Average ms per call of fn over iters, HIP events on the launch stream."""
    import torch
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(iters):
        fn()
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / iters


def cpu_baseline(n, c, h, w, k) -> dict:
    """The reference's CPU checker conv2dcpu (include/verfiy.h:5-43), compiled from its source
    into oracle/_ref when available, else the oracle's faithful C restatement; single thread,
    on one image of the workload (the cost is linear in N)."""
    import numpy as np

    from oracle import cdirect
    rng = np.random.default_rng(0)
    x = rng.standard_normal((1, c, h, w)).astype(np.float16)
    wt = (rng.standard_normal((k, c, 3, 3)) * np.sqrt(2.0 / (9 * c))).astype(np.float16)
    kind = "reference" if cdirect.ref_lib() is not None else "port"
    fn = cdirect.ref_conv2dcpu if kind == "reference" else cdirect.conv2dcpu_faithful
    t0 = time.perf_counter()
    fn(x, wt)
    dt = time.perf_counter() - t0
    flops = 2.0 * k * c * 9 * h * w
    return {"value": flops / dt / 1e12, "unit": "TFLOP/s", "cores": 1, "kind": kind,
            "sample": f"1 of {WORKLOAD['n']} images (N=1, C={c}, {h}x{w}, K={k}), single thread, "
                      f"{dt:.2f} s, {flops / dt / 1e9:.3f} GFLOP/s",
            "seconds": dt}


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "wino_gemm_n32_c256_h56_k256.md")


def measured_traffic(cfg: int):
    """HBM bytes per GEMM launch on this workload, from the committed counter profile
    (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by its copy-kernel calibration).
    None when the profile is absent or was taken for another tile configuration."""
    import re
    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return None
    m = re.search(r"^traffic_per_launch: cfg=(\d+) bytes=(\d+)", text, flags=re.M)
    if m is None or int(m.group(1)) != cfg:
        return None
    return int(m.group(2))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cfg", type=int, default=-1, help="GEMM tile configuration (-1 = library choice)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="before warm-up, run one device-to-device copy of this many MiB (untimed): the "
                         "known-size kernel that hardware byte counters are calibrated against")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported in this build "
              "(the natural multi-GPU split, by images, needs no exchange).", file=sys.stderr)
        return 2

    import torch
    runner = importlib.import_module("2024pra_conv_amd.runner")
    dev = torch.device("cuda:0")
    n, c, h, w, k = (WORKLOAD[t] for t in ("n", "c", "h", "w", "k"))
    gen = torch.Generator(device=dev)
    gen.manual_seed(20240)
    x = torch.randn((n, c, h, w), generator=gen, device=dev, dtype=torch.float32).half()
    wt = (torch.randn((k, c, 3, 3), generator=gen, device=dev, dtype=torch.float32) * (2.0 / (9 * c)) ** 0.5).half()
    conv = runner.WinogradConv(x, wt)
    if args.cfg >= 0:
        conv.L.conv2d_tuned_clear()
    cfg = conv.selected_config() if args.cfg < 0 else args.cfg

    oh, ow = conv.oh, conv.ow
    T = ((oh + 1) // 2) * ((ow + 1) // 2)
    direct_flops = 2.0 * n * k * c * 9 * oh * ow
    gemm_flops = 32.0 * n * k * c * T
    v_bytes = 32.0 * n * c * T          # algorithmic V bytes (FP16, unpadded)
    u_bytes = 32.0 * k * c
    in_bytes = 2.0 * n * c * h * w
    out_bytes = 2.0 * n * k * oh * ow

    step = (lambda: conv.launch_slots()) if args.cfg < 0 else (
        lambda: (conv.prepare_weights(), conv.input_transform(), conv.gemm(cfg)))
    if args.copy_mib > 0:
        src = torch.empty(args.copy_mib << 20, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms = _events_time(step, args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    # per-stage and steady-state timings (same stream, HIP events)
    conv.prepare_weights()
    ms_s1 = _events_time(conv.prepare_weights, args.steps)
    ms_s2 = _events_time(conv.input_transform, args.steps)
    ms_s3 = _events_time(lambda: conv.gemm(cfg), args.steps)
    ms_steady = _events_time(lambda: conv.forward(cfg), args.steps)

    # accuracy of the benchmarked output against the FP64 direct convolution
    step()
    ref = conv.direct_f64()
    torch.cuda.synchronize()
    y = conv.out.double()
    d = (y - ref).abs()
    max_rel = (d.max() / ref.abs().max()).item()
    rms_rel = (d.norm() / ref.norm()).item()
    finite = bool(torch.isfinite(y).all())

    s3_bytes = v_bytes + u_bytes + out_bytes
    s3_tflops = gemm_flops / (ms_s3 * 1e-3) / 1e12
    s3_gbs = s3_bytes / (ms_s3 * 1e-3) / 1e9
    mfma_floor = gemm_flops / (PEAK_FP16_TFLOPS * 1e12)
    hbm_floor = s3_bytes / (PEAK_HBM_GBS * 1e9)
    if hbm_floor > mfma_floor:
        roof = {"bound": "hbm", "achieved": round(s3_gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(s3_gbs / PEAK_HBM_GBS, 4)}
    else:
        roof = {"bound": "mfma", "achieved": round(s3_tflops, 1), "peak": PEAK_FP16_TFLOPS, "unit": "TFLOP/s",
                "frac": round(s3_tflops / PEAK_FP16_TFLOPS, 4)}
    roof.update({"kernel": "wino_gemm (S3, fused output transform)", "traffic": measured_traffic(cfg),
                 "mfma_frac": round(s3_tflops / PEAK_FP16_TFLOPS, 4),
                 "hbm_frac": round(s3_gbs / PEAK_HBM_GBS, 4),
                 "algorithmic_bytes": s3_bytes, "algorithmic_flops": gemm_flops})

    value = direct_flops / (ms * 1e-3) / 1e12
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16",
        "data": "synthetic: seeded N(0,1) FP16 input, N(0,2/(9C)) FP16 filters",
        "config": {"workload": f"plain FP16 Winograd F(2x2,3x3) conv2d, N={n} C={c} H=W={h} K={k}, 3x3 stride 1 pad 1, "
                               "four-slot reference protocol per step (filter transform included)",
                   "tile_config": cfg},
        "roofline": roof,
        "stages_us": {"S1_filter": round(ms_s1 * 1e3, 2), "S2_input": round(ms_s2 * 1e3, 2),
                      "S3_gemm_fused": round(ms_s3 * 1e3, 2), "steady_S2_S3": round(ms_steady * 1e3, 2)},
        "steady_state_tflops": round(direct_flops / (ms_steady * 1e-3) / 1e12, 2),
        "s2_hbm_frac": round((in_bytes + v_bytes) / (ms_s2 * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
        "error_vs_fp64": {"max_rel": max_rel, "rms_rel": rms_rel, "finite": finite,
                          "definition": "max|y-y*|/max|y*| and ||y-y*||/||y*|| vs FP64 direct conv"},
        "wall_s_timed": round(wall, 4),
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(1, c, h, w, k)
        line["cpu_baseline"].pop("seconds", None)
        line["cpu_baseline"]["value"] = round(line["cpu_baseline"]["value"], 6)
    ok = finite and max_rel <= 1e-2
    print(json.dumps(line), flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
