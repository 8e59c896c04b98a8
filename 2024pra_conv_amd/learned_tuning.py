"""Learned tuning: collect autotuner timings, train the MLP cost model on the GPU (the
library's HIP kernels), and measure its regret against exhaustive autotuning.

Workload (the learned-tuning configuration): shapes drawn as in the 24-layer batched
workload (N in {1,4,16}, C,K in {3,32,64,96,128,256,512}, H=W in {7,13,14,27,28,56,112},
3x3 stride 1 pad 1), 2000 training shapes x 16 tile configurations, 200 held-out shapes.

Two choices are learned, each by networks of the same 21 -> 64 -> 64 -> 1 architecture:
  * the GEMM tile configuration (one network on GEMM times, conv2d_mlp_select);
  * the form and configuration of the whole forward pass (conv2d_mlp_select_form): one network
    on the staged form's time (input transform + GEMM) per configuration, one on the fused
    kernel's time; the choice is the smallest prediction among the 16 staged options and the
    fused form.

  python 2024pra_conv_amd/learned_tuning.py --train 2000 --test 200   (from the repo root)

Prints one JSON line: mean / median / p90 regret against the exhaustive best of the learned
choice, of the library's hand rules (the occupancy heuristic for the configuration and the
measured-threshold rule for the form), of a static K-threshold dispatch and of a random choice,
and the tuning wall-clock of exhaustive search vs. MLP inference. Saves the trained parameters
(.npy) and the measurement table (.npz) when --out is given.
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
    """Exhaustive autotuner timings (us) for each shape: GEMM time of every configuration (S, ncfg),
    the input transform (S,) and the fused kernel (S,; -1 where it does not apply), and the
    library's own choices before tuning: configuration (S,) and form (S,)."""
    L = wc.load_library()
    dev = torch.device("cuda:0")
    ncfg = L.conv2d_num_tile_configs()
    out = np.zeros((len(shapes), ncfg), dtype=np.float64)
    s2 = np.zeros(len(shapes), dtype=np.float64)
    fused = np.zeros(len(shapes), dtype=np.float64)
    heur = np.zeros(len(shapes), dtype=np.int64)
    form = np.zeros(len(shapes), dtype=np.int64)
    t0 = time.perf_counter()
    for i, (n, c, h, k) in enumerate(shapes):
        x = torch.empty((n, c, h, h), dtype=torch.float16, device=dev).uniform_(-1, 1)
        w = torch.empty((k, c, 3, 3), dtype=torch.float16, device=dev).uniform_(-1, 1)
        conv = runner.WinogradConv(x, w)
        L.conv2d_tuned_clear()
        heur[i] = conv.selected_config()
        form[i] = conv.form()
        conv.prepare_weights()
        conv.input_transform()
        _, times, s2[i], fused[i] = conv.autotune(reps=reps, forms=True)
        out[i] = times
        if log_every and (i + 1) % log_every == 0:
            print(f"measured {i + 1}/{len(shapes)} shapes, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    L.conv2d_tuned_clear()
    return {"gemm": out, "s2": s2, "fused": fused, "heur_cfg": heur, "heur_form": form,
            "seconds": time.perf_counter() - t0}


def option_times(m) -> np.ndarray:
    """Whole-forward time of every option per shape: the 16 staged configurations (input
    transform + GEMM) then the fused form (inf where it does not apply): (S, ncfg + 1)."""
    staged = m["gemm"] + m["s2"][:, None]
    fz = np.where(m["fused"] > 0, m["fused"], np.inf)
    return np.concatenate([staged, fz[:, None]], axis=1)


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
        if done % 1000 == 0 or done == steps:
            print(f"trained {done}/{steps} steps on {X.shape[0]} rows, loss {loss.value:.4f}, "
                  f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
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


def select_form(Ps: torch.Tensor, Pf: torch.Tensor, shapes, ncfg: int):
    """Learned option per shape: the configuration index for the staged form, ncfg for fused."""
    L = wc.load_library()
    best, form = ctypes.c_int(), ctypes.c_int()
    out = []
    t0 = time.perf_counter()
    for (n, c, h, k) in shapes:
        pb = wc.make_problem(n, c, h, h, k)
        wc.check(L.conv2d_mlp_select_form(_ptr(Ps), _ptr(Pf), ctypes.byref(pb), ctypes.byref(form), ctypes.byref(best),
                                          None, _stream()), "conv2d_mlp_select_form")
        out.append(ncfg if form.value == wc.FORM_FUSED else best.value)
    return np.array(out), time.perf_counter() - t0


def train_form_models(shapes, m, steps: int, lr: float, seed: int):
    """The two networks of the form choice, trained on measured whole-forward times."""
    ncfg = m["gemm"].shape[1]
    X = features(shapes, ncfg)
    Ps, loss_s, ts = train(X, np.log((m["gemm"] + m["s2"][:, None]).reshape(-1)), steps, lr, seed)
    ok = m["fused"] > 0
    Xf = X.reshape(len(shapes), ncfg, -1)[ok, 0, :]
    Pf, loss_f, tf = train(np.ascontiguousarray(Xf), np.log(m["fused"][ok]), steps, lr, seed + 1)
    return Ps, Pf, loss_s, loss_f, ts + tf


def static_k_dispatch(shapes) -> np.ndarray:
    """A static tile choice by output-channel count alone (the style of a hand-written
    per-shape dispatch): 128x128 for K >= 128, 64x128 for K >= 64, 32x128 for K >= 32."""
    return np.array([0 if k >= 128 else (1 if k >= 64 else (4 if k >= 32 else 6)) for (_, _, _, k) in shapes])


def regret(times: np.ndarray, choice: np.ndarray) -> np.ndarray:
    best = times.min(axis=1)
    return times[np.arange(len(choice)), choice] / best - 1.0


def stats(r):
    return {"mean_pct": round(100 * float(r.mean()), 2), "median_pct": round(100 * float(np.median(r)), 2),
            "p90_pct": round(100 * float(np.quantile(r, 0.9)), 2), "exact_best_frac": round(float(np.mean(r == 0)), 3)}


def run(n_train: int, n_test: int, steps: int, lr: float, reps: int, seed: int, out: str | None) -> dict:
    ncfg = wc.load_library().conv2d_num_tile_configs()
    tr_shapes = draw_shapes(seed, n_train)
    te_shapes = draw_shapes(seed + 1, n_test)
    tr = measure(tr_shapes, reps)
    te = measure(te_shapes, reps)
    rng = np.random.default_rng(seed + 2)
    # (1) GEMM tile configuration (S3 times only)
    X = features(tr_shapes, ncfg)
    P, final_loss, train_s = train(X, np.log(tr["gemm"].reshape(-1)), steps, lr, seed)
    mlp_choice, infer_s = select(P, te_shapes)
    te_t = te["gemm"]
    gemm = {"regret_mlp": stats(regret(te_t, mlp_choice)), "regret_library_heuristic": stats(regret(te_t, te["heur_cfg"])),
            "regret_static_k_dispatch": stats(regret(te_t, static_k_dispatch(te_shapes))),
            "regret_random": stats(regret(te_t, rng.integers(0, ncfg, len(te_shapes)))),
            "final_train_mse_log_time": round(final_loss, 5), "train_seconds": round(train_s, 2)}
    # (2) form + configuration of the whole forward pass
    Ps, Pf, loss_s, loss_f, train_f = train_form_models(tr_shapes, tr, steps, lr, seed)
    opt = option_times(te)
    f_choice, f_infer = select_form(Ps, Pf, te_shapes, ncfg)
    lib_choice = np.where(te["heur_form"] == wc.FORM_FUSED, ncfg, te["heur_cfg"])
    r_form_mlp, r_form_lib = regret(opt, f_choice), regret(opt, lib_choice)
    form = {"options": f"{ncfg} staged configurations (input transform + GEMM) + the fused form",
            "regret_mlp": stats(r_form_mlp), "regret_library_rules": stats(r_form_lib),
            "regret_staged_static_k_dispatch": stats(regret(opt, static_k_dispatch(te_shapes))),
            "regret_random": stats(regret(opt, rng.integers(0, ncfg + 1, len(te_shapes)))),
            "fused_best_frac": round(float(np.mean(opt.argmin(axis=1) == ncfg)), 3),
            "mlp_form_agrees_with_best": round(float(np.mean((f_choice == ncfg) == (opt.argmin(axis=1) == ncfg))), 3),
            "library_form_agrees_with_best": round(float(np.mean((lib_choice == ncfg) == (opt.argmin(axis=1) == ncfg))), 3),
            "mlp_gain_over_library_rules_pct": round(100 * float(r_form_lib.mean() - r_form_mlp.mean()), 2),
            "held_out_total_time_us": {"exhaustive_best": round(float(opt.min(axis=1).sum()), 1),
                                       "mlp_choice": round(float(opt[np.arange(n_test), f_choice].sum()), 1),
                                       "library_rules": round(float(opt[np.arange(n_test), lib_choice].sum()), 1)},
            "final_train_mse_log_time": {"staged": round(loss_s, 5), "fused": round(loss_f, 5)},
            "train_seconds": round(train_f, 2)}
    if out:
        os.makedirs(out, exist_ok=True)
        np.save(os.path.join(out, "mlp_params.npy"), P.cpu().numpy())
        np.save(os.path.join(out, "mlp_staged_params.npy"), Ps.cpu().numpy())
        np.save(os.path.join(out, "mlp_fused_params.npy"), Pf.cpu().numpy())
        np.savez_compressed(os.path.join(out, "timings.npz"), train_shapes=np.array(tr_shapes),
                            test_shapes=np.array(te_shapes), **{f"train_{k}": v for k, v in tr.items() if k != "seconds"},
                            **{f"test_{k}": v for k, v in te.items() if k != "seconds"})
    return {
        "workload": f"learned tuning: {n_train} train / {n_test} held-out shapes x {ncfg} tile configs + fused form",
        "mlp": "21-64-64-1 FP32, full-batch Adam on GPU (library HIP kernels)",
        "train_steps": steps,
        "gemm_config": gemm,
        "form_and_config": form,
        "tuning_wall_s": {"exhaustive_per_shape": round(te["seconds"] / n_test, 4),
                          "mlp_inference_per_shape": round(f_infer / n_test, 5),
                          "saved_frac": round(1.0 - f_infer / te["seconds"], 4)},
        "measurement_wall_s": round(tr["seconds"] + te["seconds"], 1),
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
