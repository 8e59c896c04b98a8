"""GPU tests of the persisted transformed weights (conv2d_weights_save / _load): a round
trip restores U bit for bit on the Winograd and the direct path, and a file for other
weights or with a corrupted payload is refused and leaves U untouched."""
import numpy as np
import pytest
import torch

from oracle import winograd as wg

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("geom", [(2, 40, 14, 22, 70, 3, 1, 1), (1, 8, 11, 13, 33, 5, 2, 2)], ids=["winograd", "direct"])
def test_round_trip_and_refusals(geom, runner, wc, cuda, tmp_path):
    n, c, h, w, k, r, st, p = geom
    print(f"case {geom}", flush=True)
    x = torch.from_numpy(wg.uniform_pm1(40 + c, (n, c, h, w))).to(cuda)
    wt = torch.from_numpy(wg.uniform_pm1(41 + c, (k, c, r, r))).to(cuda)
    a = runner.WinogradConv(x, wt, pad=(p, p), stride=(st, st))
    a.prepare_weights()
    a.forward()
    torch.cuda.synchronize()
    y_ref = a.out.clone()
    path = str(tmp_path / "u.bin")
    a.save_weights(path)

    b = runner.WinogradConv(x, wt.clone(), pad=(p, p), stride=(st, st))
    b.U.fill_(0x55)
    assert b.load_weights(path) == wc.OK
    # the U of the problem's form (the buffer is sized for every form)
    assert torch.equal(b.U[: b.u_bytes], a.U[: a.u_bytes])
    assert bool((b.U[b.u_bytes:] == 0x55).all())
    b.forward()
    torch.cuda.synchronize()
    assert torch.equal(b.out.view(torch.int16), y_ref.view(torch.int16))

    # other weights: refused, U untouched
    w2 = wt.clone()
    w2.view(-1)[7] += 1.0
    c2 = runner.WinogradConv(x, w2, pad=(p, p), stride=(st, st))
    c2.U.fill_(0x55)
    assert c2.load_weights(path) == wc.ERR_MISMATCH
    assert bool((c2.U == 0x55).all())

    # corrupted payload: refused
    data = bytearray(open(path, "rb").read())
    data[64 + 100] ^= 0x01
    bad = str(tmp_path / "bad.bin")
    open(bad, "wb").write(bytes(data))
    b.U.fill_(0x55)
    assert b.load_weights(bad) == wc.ERR_MISMATCH
    assert bool((b.U == 0x55).all())


def test_round_trip_of_the_f43_form(runner, wc, cuda, tmp_path):
    """A shape tuned to the F(4x4) form saves and loads that form's U (its own header layout);
    a file written for the F(2x2) form of the same weights is refused."""
    n, c, h, w, k = 2, 40, 14, 22, 70
    x = torch.from_numpy(wg.uniform_pm1(60, (n, c, h, w))).to(cuda)
    wt = torch.from_numpy(wg.uniform_pm1(61, (k, c, 3, 3))).to(cuda)
    L = wc.load_library()
    L.conv2d_tuned_clear()
    try:
        a = runner.WinogradConv(x, wt)
        a.prepare_weights()
        f23 = str(tmp_path / "u23.bin")
        a.save_weights(f23)
        a.set_form(wc.FORM_F43, 1)
        a.prepare_weights()
        a.forward()
        torch.cuda.synchronize()
        y_ref = a.out.clone()
        path = str(tmp_path / "u43.bin")
        a.save_weights(path)
        b = runner.WinogradConv(x, wt.clone())
        b.U.fill_(0x55)
        assert b.load_weights(f23) == wc.ERR_MISMATCH and bool((b.U == 0x55).all())
        assert b.load_weights(path) == wc.OK
        assert torch.equal(b.U[: b.u43_bytes], a.U[: a.u43_bytes])
        b.forward()
        torch.cuda.synchronize()
        assert torch.equal(b.out.view(torch.int16), y_ref.view(torch.int16))
    finally:
        L.conv2d_tuned_clear()
