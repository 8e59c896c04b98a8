"""Seeded random geometries through every form the library offers for them, against the GPU FP64
direct convolution (conv2d_direct_f64, itself pinned bit for bit to the CPU oracle).

Widths, heights, channel counts and pads are drawn so that odd and one-pixel widths, channel
counts that are not multiples of 8, 16 or 32 and output-channel counts whose last 64-channel
block is half empty all occur; a fifth of the shapes take other filter sizes and strides (the
direct path). Bar: normwise max relative error <= 1e-2 and finite outputs, for the default
form (the four slots getkernelInfo fills) and, on the Winograd path, for the fused form and the
staged form (input transform + GEMM) where each applies.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _shapes(count=36, seed=2026):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        n = int(rng.integers(1, 4))
        c = int(rng.choice([1, 3, 5, 8, 13, 16, 24, 31, 33, 40, 64, 70, 96, 130]))
        k = int(rng.choice([1, 3, 7, 16, 31, 32, 33, 64, 65, 96, 100, 160]))
        h = int(rng.integers(1, 48))
        w = int(rng.choice([1, 2, 3, 5, 7, 9, 14, 27, 28, 31, 33, 55, 56, 63]))
        if rng.random() < 0.2:  # direct path
            r = s = int(rng.choice([1, 3, 5]))
            u = v = int(rng.choice([1, 2]))
            p = q = int(rng.integers(0, r // 2 + 1))
        else:
            r = s = 3
            u = v = 1
            p = int(rng.integers(0, 3))
            q = int(rng.integers(0, 3))
        if (h + 2 * p - r) // u + 1 < 1 or (w + 2 * q - s) // v + 1 < 1:
            continue
        out.append((n, c, h, w, k, r, s, u, v, p, q))
    return out


SHAPES = _shapes()


def _ids(s):
    return "x".join(map(str, s))


@pytest.mark.parametrize("shape", SHAPES, ids=_ids)
def test_random_geometry_every_form(shape, runner, wc, cuda):
    n, c, h, w, k, r, s, u, v, p, q = shape
    print(f"case {shape}", flush=True)
    gen = torch.Generator(device=cuda)
    gen.manual_seed(hash(shape) & 0xFFFFFFFF)
    x = (torch.rand((n, c, h, w), generator=gen, device=cuda) * 2 - 1).half()
    wt = (torch.rand((k, c, r, s), generator=gen, device=cuda) * 2 - 1).half()
    conv = runner.WinogradConv(x, wt, pad=(p, q), stride=(u, v))
    conv.L.conv2d_tuned_clear()
    ref = conv.direct_f64()
    scale = ref.abs().max().item()

    def check(what):
        torch.cuda.synchronize()
        y = conv.out.double()
        assert bool(torch.isfinite(y).all()), f"{what}: non-finite output"
        rel = (y - ref).abs().max().item() / scale if scale > 0 else (y - ref).abs().max().item()
        print(f"  {what}: max_rel {rel:.3e}", flush=True)
        assert rel <= 1e-2, f"{what}: max_rel {rel}"

    # default form through the reference's four-slot protocol
    conv.prepare_weights()
    conv.out.fill_(float("nan"))
    conv.launch_slots()
    check(f"four slots (method {conv.method()}, form {conv.form()})")
    if conv.method() != wc.METHOD_WINOGRAD:
        return
    # fused and staged forms where they apply (UNSUPPORTED otherwise)
    for what, run in (("fused", lambda: conv.fused_gemm()),
                      ("staged", lambda: (conv.input_transform(), conv.gemm()))):
        conv.out.fill_(float("nan"))
        try:
            run()
        except wc.Conv2dError as e:
            assert e.code == wc.ERR_UNSUPPORTED, e
            print(f"  {what}: not offered for this geometry", flush=True)
            continue
        check(what)
