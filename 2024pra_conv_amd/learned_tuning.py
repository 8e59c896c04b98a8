"""Learned tile-configuration tuning: collect autotuner timings, train the MLP cost model on
the GPU (the library's HIP kernels), and measure its regret against exhaustive autotuning.

Workload (the learned-tuning configuration): shapes drawn as in the 24-layer batched
workload (N in {1,4,16}, C,K in {3,32,64,96,128,256,512}, H=W in {7,13,14,27,28,56,112},
3x3 stride 1 pad 1), 2000 training shapes x 16 tile configurations, 200 held-out shapes.

  python -m 2024pra_conv_amd.learned_tuning --train 2000 --test 200   (from the repo root,
  via: python 2024pra_conv_amd/learned_tuning.py ...)

Prints one JSON line: mean / median / p90 regret of the MLP's choice, of the library's shape
heuristic, of a static K-threshold dispatch and of a random choice against the exhaustive
best, and the tuning
wall-clock of exhaustive search vs. MLP inference. Saves the trained parameters (.npy) and
the measurement table (.npz) when --out is given, so tuning and training can resume.
"""
from __future__ import annotations

import argparse
import ctypes
import importlib
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

wc = importlib.import_module("2024pra_conv_amd")
runner = importlib.import_module("2024pra_conv_amd.runner")

NS, CS, HS = [1, 4, 16], [3, 32, 64, 96, 128, 256, 512], [7, 13, 14, 27, 28, 56, 112]


def draw_shapes(seed: int, count: int):
    rng = np.random.default_rng(seed)
    return [(int(rng.choice(NS)), int(rng.choice(CS)), int(rng.choice(HS)), int(rng.choice(CS)))
            for _ in range(count)]


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def measure(shapes, reps: int = 2, log_every: int = 100):
    """Exhaustive autotuner timings (us) of every configuration for each shape: (S, ncfg)."""
    L = wc.load_library()
    dev = torch.device("cuda:0")
    ncfg = L.conv2d_num_tile_configs()
    out = np.zeros((len(shapes), ncfg), dtype=np.float64)
    heur = np.zeros(len(shapes), dtype=np.int64)
    t0 = time.perf_counter()
    for i, (n, c, h, k) in enumerate(shapes):
        x = torch.empty((n, c, h, h), dtype=torch.float16, device=dev).uniform_(-1, 1)
        w = torch.empty((k, c, 3, 3), dtype=torch.float16, device=dev).uniform_(-1, 1)
        conv = runner.WinogradConv(x, w)
        L.conv2d_tuned_clear()
        heur[i] = conv.selected_config()
        conv.prepare_weights()
        conv.input_transform()
        _, times = conv.autotune(reps=reps)
        out[i] = times
        if log_every and (i + 1) % log_every == 0:
            print(f"measured {i + 1}/{len(shapes)} shapes, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    L.conv2d_tuned_clear()
    return out, heur, time.perf_counter() - t0


def features(shapes, ncfg):
    L = wc.load_library()
    X = np.zeros((len(shapes) * ncfg, 21), dtype=np.float32)
    buf = (ctypes.c_float * 21)()
    for i, (n, c, h, k) in enumerate(shapes):
        pb = wc.make_problem(n, c, h, h, k)
        for cfg in range(ncfg):
            wc.check(L.conv2d_mlp_features(ctypes.byref(pb), cfg, buf), "conv2d_mlp_features")
            X[i * ncfg + cfg] = np.frombuffer(buf, dtype=np.float32)
    return X


def init_params(seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, shape in ((21, (64, 21)), (None, (64,)), (64, (64, 64)), (None, (64,)), (64, (64,)), (None, (1,))):
        parts.append(np.zeros(shape) if fan_in is None else rng.uniform(-1, 1, shape) * np.sqrt(6.0 / fan_in))
    return np.concatenate([a.ravel() for a in parts]).astype(np.float32)


def train(X: np.ndarray, T: np.ndarray, steps: int, lr: float, seed: int = 0):
    """Full-batch Adam on the GPU with the library's kernels. Returns (params, final loss, seconds)."""
    L = wc.load_library()
    dev = torch.device("cuda:0")
    P = torch.from_numpy(init_params(seed)).to(dev)
    M, V = torch.zeros_like(P), torch.zeros_like(P)
    Xd, Td = torch.from_numpy(X).to(dev), torch.from_numpy(T.astype(np.float32)).to(dev)
    loss = ctypes.c_float()
    t0 = time.perf_counter()
    done = 0
    chunk = 250
    while done < steps:
        s = min(chunk, steps - done)
        wc.check(L.conv2d_mlp_train(_ptr(P), _ptr(M), _ptr(V), done, _ptr(Xd), _ptr(Td), X.shape[0], s,
                                    ctypes.c_float(lr), ctypes.byref(loss), _stream()), "conv2d_mlp_train")
        done += s
    torch.cuda.synchronize()
    return P, loss.value, time.perf_counter() - t0


def select(P: torch.Tensor, shapes):
    L = wc.load_library()
    best = ctypes.c_int()
    out = []
    t0 = time.perf_counter()
    for (n, c, h, k) in shapes:
        pb = wc.make_problem(n, c, h, h, k)
        wc.check(L.conv2d_mlp_select(_ptr(P), ctypes.byref(pb), ctypes.byref(best), None, _stream()), "conv2d_mlp_select")
        out.append(best.value)
    return np.array(out), time.perf_counter() - t0


def static_k_dispatch(shapes) -> np.ndarray:
    """A static tile choice by output-channel count alone (the style of a hand-written
    per-shape dispatch): 128x128 for K >= 128, 64x128 for K >= 64, 32x128 for K >= 32."""
    return np.array([0 if k >= 128 else (1 if k >= 64 else (4 if k >= 32 else 6)) for (_, _, _, k) in shapes])


def regret(times: np.ndarray, choice: np.ndarray) -> np.ndarray:
    best = times.min(axis=1)
    return times[np.arange(len(choice)), choice] / best - 1.0


def run(n_train: int, n_test: int, steps: int, lr: float, reps: int, seed: int, out: str | None) -> dict:
    ncfg = wc.load_library().conv2d_num_tile_configs()
    tr_shapes = draw_shapes(seed, n_train)
    te_shapes = draw_shapes(seed + 1, n_test)
    tr_times, _, tr_wall = measure(tr_shapes, reps)
    te_times, te_heur, te_wall = measure(te_shapes, reps)
    X = features(tr_shapes, ncfg)
    T = np.log(tr_times.reshape(-1))
    P, final_loss, train_s = train(X, T, steps, lr, seed)
    mlp_choice, infer_s = select(P, te_shapes)
    r_mlp = regret(te_times, mlp_choice)
    r_heur = regret(te_times, te_heur)
    r_static = regret(te_times, static_k_dispatch(te_shapes))
    rng = np.random.default_rng(seed + 2)
    r_rand = regret(te_times, rng.integers(0, ncfg, len(te_shapes)))
    if out:
        os.makedirs(out, exist_ok=True)
        np.save(os.path.join(out, "mlp_params.npy"), P.cpu().numpy())
        np.savez_compressed(os.path.join(out, "timings.npz"), train_shapes=np.array(tr_shapes),
                            train_times=tr_times, test_shapes=np.array(te_shapes), test_times=te_times)

    def stats(r):
        return {"mean_pct": round(100 * float(r.mean()), 2), "median_pct": round(100 * float(np.median(r)), 2),
                "p90_pct": round(100 * float(np.quantile(r, 0.9)), 2), "exact_best_frac": round(float(np.mean(r == 0)), 3)}

    return {
        "workload": f"learned tuning: {n_train} train / {n_test} held-out shapes x {ncfg} tile configs",
        "mlp": "21-64-64-1 FP32, full-batch Adam on GPU (library HIP kernels)",
        "train_steps": steps, "final_train_mse_log_time": round(final_loss, 5), "train_seconds": round(train_s, 2),
        "regret_mlp": stats(r_mlp), "regret_library_heuristic": stats(r_heur),
        "regret_static_k_dispatch": stats(r_static), "regret_random": stats(r_rand),
        "held_out_total_time_us": {"exhaustive_best": round(float(te_times.min(axis=1).sum()), 1),
                                   "mlp_choice": round(float(te_times[np.arange(n_test), mlp_choice].sum()), 1),
                                   "heuristic": round(float(te_times[np.arange(n_test), te_heur].sum()), 1)},
        "tuning_wall_s": {"exhaustive_per_shape": round(te_wall / n_test, 4),
                          "mlp_inference_per_shape": round(infer_s / n_test, 5),
                          "saved_frac": round(1.0 - infer_s / te_wall, 4)},
        "measurement_wall_s": round(tr_wall + te_wall, 1),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--test", type=int, default=200)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    print(json.dumps(run(a.train, a.test, a.steps, a.lr, a.reps, a.seed, a.out)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
