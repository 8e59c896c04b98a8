"""GPU tests of the workload configurations at their stated sizes, through the C-ABI.

  cfg1  N=1 C=64 56x56 K=64, uniform[-1,1]: the library's default form (the four-slot protocol
        getkernelInfo fills) against its oracle model and the FP64 direct convolution.
  cfg4  the 24 seeded layers of the batched workload: batched == separate launches bit for bit,
        graph replays == eager, every layer within max_rel 1e-2 of the GPU FP64 kernel, before
        and after the batch autotuner.
  cfg5  learned tuning at 2000 training / 200 held-out shapes: regret of the learned choices
        against exhaustive autotuning, next to the library's hand rules.
"""
import importlib

import numpy as np
import pytest
import torch

from oracle import cdirect
from oracle import winograd as wg

pytestmark = pytest.mark.gpu


def _to_dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def test_cfg1_default_form_against_oracle(runner, wc, cuda):
    n, c, h, w, k = 1, 64, 56, 56, 64
    print("case cfg1 default form", flush=True)
    geo = wg.Geometry(n, c, h, w, k)
    x = wg.uniform_pm1(1001, (n, c, h, w))
    wt = wg.uniform_pm1(1002, (k, c, 3, 3))
    conv = runner.WinogradConv(_to_dev(x, cuda), _to_dev(wt, cuda))
    conv.L.conv2d_tuned_clear()
    assert conv.form() == wc.FORM_FUSED  # K = 64 <= 384
    conv.out.fill_(float("nan"))
    conv.launch_slots()
    torch.cuda.synchronize()
    y = conv.out.cpu().numpy().astype(np.float64)
    # U bit-exact against the oracle's filter transform
    u = conv.u_halves().cpu().numpy()
    assert np.array_equal(u.view(np.uint16), wg.pack_U(wg.filter_transform_fp32(wt), geo).view(np.uint16))
    # within the FP32-accumulation bound of the oracle's model of the fused kernel
    y_model, scale = wg.fused_gpu_model(x, wt, geo)
    e = (c + 16) * 2.0 ** -24 * scale
    tol = 2.0 ** -11 * (np.abs(y_model) + e) + e + 2.0 ** -25
    assert np.all(np.abs(y - y_model) <= tol), f"max excess {np.max(np.abs(y - y_model) - tol)}"
    ref = cdirect.conv2d_f64(x, wt)
    m = wg.error_metrics(y, ref)
    print(f"  max_rel={m['max_rel']:.3e} rms_rel={m['rms_rel']:.3e}", flush=True)
    assert m["finite"] and m["max_rel"] <= 1e-2


def _check_batch(b, convs, refs, s, what):
    for cv in convs:
        cv.out.fill_(float("nan"))
    torch.cuda.synchronize()
    b.forward(s)
    s.synchronize()
    for i, cv in enumerate(convs):
        assert torch.equal(cv.out.view(torch.int16), refs[i].view(torch.int16)), f"{what}: layer {i}"


def test_cfg4_24_layers_batched_graph_and_fp64(runner, wc, cuda):
    shapes = runner.batch_layer_shapes()
    assert len(shapes) == 24
    print(f"case cfg4 {shapes}", flush=True)
    gen = torch.Generator(device=cuda)
    gen.manual_seed(20242)
    convs = []
    for (n, c, h, k) in shapes:
        x = (torch.rand((n, c, h, h), generator=gen, device=cuda) * 2 - 1).half()
        wt = (torch.rand((k, c, 3, 3), generator=gen, device=cuda) * 2 - 1).half()
        convs.append(runner.WinogradConv(x, wt))
    conv_lib = convs[0].L
    conv_lib.conv2d_tuned_clear()
    b = runner.Batch(convs)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        b.prepare_weights(s)
        for phase in ("default", "autotuned"):
            if phase == "autotuned":
                b.autotune(reps=2, stream=s)
            print(f"  {phase}: forms {b.layer_forms()} configs {b.layer_configs()}", flush=True)
            b.forward_separate(s)
            s.synchronize()
            refs = [cv.out.clone() for cv in convs]
            # every layer against the FP64 direct convolution on the GPU
            worst = 0.0
            for i, cv in enumerate(convs):
                y64 = cv.direct_f64(s)
                s.synchronize()
                d = (refs[i].double() - y64).abs()
                rel = (d.max() / y64.abs().max()).item()
                worst = max(worst, rel)
                assert bool(torch.isfinite(refs[i]).all()) and rel <= 1e-2, f"{phase}: layer {i} max_rel {rel}"
            print(f"  {phase}: worst layer max_rel {worst:.3e}", flush=True)
            # the multi-layer launches equal the separate per-layer launches bit for bit
            _check_batch(b, convs, refs, s, f"{phase} eager")
            # one captured graph, replayed with eager work between replays
            b.capture_graph(s)
            for rep in range(3):
                for cv in convs:
                    cv.out.fill_(float("nan"))
                s.synchronize()
                b.launch_graph(s)
                s.synchronize()
                for i, cv in enumerate(convs):
                    assert torch.equal(cv.out.view(torch.int16), refs[i].view(torch.int16)), \
                        f"{phase}: replay {rep} layer {i}"
                _ = torch.ones(1000, device=cuda).sum().item()
    b.close()


@pytest.mark.timeout(1500)
def test_cfg5_learned_tuning_2000_200(wc, cuda):
    """Autotuner timings of 2000 seeded shapes train the networks on the GPU; on 200 held-out
    shapes their choices are compared with exhaustive search, with the library's hand rules and
    with a static K-threshold dispatch, for the GEMM configuration alone and for the form plus
    configuration of the whole forward pass."""
    lt = importlib.import_module("2024pra_conv_amd.learned_tuning")
    r = lt.run(2000, 200, steps=3000, lr=2e-3, reps=2, seed=11, out=None)
    g, f = r["gemm_config"], r["form_and_config"]
    print(f"gemm: mlp {g['regret_mlp']} heuristic {g['regret_library_heuristic']} "
          f"static {g['regret_static_k_dispatch']}", flush=True)
    print(f"form: mlp {f['regret_mlp']} rules {f['regret_library_rules']} "
          f"static {f['regret_staged_static_k_dispatch']} fused best {f['fused_best_frac']} "
          f"agree mlp {f['mlp_form_agrees_with_best']} rules {f['library_form_agrees_with_best']}", flush=True)
    print(f"tuning {r['tuning_wall_s']} measurement {r['measurement_wall_s']} s", flush=True)
    assert np.isfinite(g["final_train_mse_log_time"])
    assert g["regret_mlp"]["mean_pct"] < g["regret_static_k_dispatch"]["mean_pct"]
    assert g["regret_mlp"]["mean_pct"] < 5.0
    assert f["regret_mlp"]["mean_pct"] < f["regret_staged_static_k_dispatch"]["mean_pct"]
    assert f["regret_mlp"]["mean_pct"] < 5.0
    assert r["tuning_wall_s"]["saved_frac"] > 0.9
