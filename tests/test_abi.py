"""C-ABI boundary tests (CPU only: no call here launches a kernel)."""
import ctypes
import heapq
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "conv2d.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|size_t|const char\*)\s+(\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol(wc):
    L = wc.load_library()
    names = _declared_functions()
    assert "getParamsize" in names and "getkernelInfo" in names and len(names) >= 18
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", wc.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (\w+)$", out, flags=re.M))
    assert set(names) <= exported
    # the bindings cover exactly the declared API
    assert set(wc.exported_symbols()) == set(names)


def test_struct_layouts_match_c_headers(wc, tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(r'''
#include <stddef.h>
#include <stdio.h>
#include "mykernelParamType.h"
int main(void) {
  printf("%zu %zu %zu\n", sizeof(problem_t), sizeof(kernelInfo_t), sizeof(mykernelParamType));
  printf("%zu %zu %zu %zu\n", offsetof(problem_t, n), offsetof(problem_t, q), offsetof(kernelInfo_t, kernelPtr1), offsetof(kernelInfo_t, blockx4));
  printf("%zu %zu %zu %zu %zu\n", offsetof(mykernelParamType, n), offsetof(mykernelParamType, Ow),
         offsetof(mykernelParamType, tiles_h), offsetof(mykernelParamType, cfg), offsetof(mykernelParamType, flags));
  printf("%zu %zu %zu %zu %zu %zu\n", offsetof(mykernelParamType, nsteps), offsetof(mykernelParamType, fnu),
         offsetof(mykernelParamType, act), offsetof(mykernelParamType, form), offsetof(mykernelParamType, bias),
         offsetof(mykernelParamType, io));
  return 0; }
''')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-Wall", "-Werror", f"-I{ROOT}/include", "-o", str(exe), str(src)], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    assert list(map(int, lines[0].split())) == [ctypes.sizeof(wc.Problem), ctypes.sizeof(wc.KernelInfo),
                                                ctypes.sizeof(wc.ParamType)]
    P, K, M = wc.Problem, wc.KernelInfo, wc.ParamType
    assert list(map(int, lines[1].split())) == [P.n.offset, P.q.offset, K.kernelPtr1.offset, K.blockx4.offset]
    assert list(map(int, lines[2].split())) == [M.n.offset, M.Ow.offset, M.tiles_h.offset, M.cfg.offset, M.flags.offset]
    assert list(map(int, lines[3].split())) == [M.nsteps.offset, M.fnu.offset, M.act.offset, M.form.offset,
                                                M.bias.offset, M.io.offset]
    # the reference layouts (probed: problem_t 96 B, kernelInfo_t 160 B; param: pointers 0-47,
    # the 13 shorts at 48-73)
    assert ctypes.sizeof(P) == 96 and ctypes.sizeof(K) == 160
    assert M.n.offset == 48 and M.Ow.offset == 72


def test_getparamsize(wc):
    L = wc.load_library()
    pb = wc.make_problem(1, 64, 56, 56, 64)
    sz = ctypes.c_int()
    assert L.getParamsize(ctypes.byref(pb), ctypes.byref(sz)) == wc.OK
    assert sz.value == ctypes.sizeof(wc.ParamType)
    assert L.getParamsize(ctypes.byref(pb), None) == wc.ERR_NULL


def _ki(wc, pb):
    L = wc.load_library()
    ki = wc.KernelInfo()
    buf = ctypes.create_string_buffer(ctypes.sizeof(wc.ParamType))
    rc = L.getkernelInfo(ctypes.byref(pb), ctypes.byref(ki), buf)
    return rc, ki, wc.ParamType.from_buffer_copy(buf)


FAKE = {"in_": 256, "weight": 256, "out": 256, "transform_pin": 256, "transform_filter": 256}


def _list_schedule(n1, d1, n2, d2, units=256):
    """Finish time of n1 workgroups of duration d1 then n2 of duration d2, each started on the
    earliest free of `units` compute units (one heap entry per workgroup: small cases only)."""
    free = [0.0] * units
    end = 0.0
    for n, d in ((n1, d1), (n2, d2)):
        for _ in range(n):
            t = heapq.heappop(free) + d
            end = max(end, t)
            heapq.heappush(free, t)
    return end


def test_list_schedule_restatement():
    # closed forms: whole rounds of equal workgroups, and a half round of longer ones after them
    assert _list_schedule(512, 1.0, 0, 1.0) == 2.0 and _list_schedule(513, 1.0, 0, 1.0) == 3.0
    assert _list_schedule(448, 1.0, 224, 1.5) == 3.5 and _list_schedule(0, 1.0, 384, 1.5) == 3.0


def test_wide_tail_items_where_the_schedule_gains(wc):
    """Fused layers whose last 64-channel block holds 1..32 channels: wide tail items (grid
    smaller than one item per 64-channel block and 64-tile block) for the measured gains, half-block
    items for the measured losses."""
    cases = [((32, 96, 55, 55, 96), True), ((32, 96, 54, 54, 96), True), ((32, 32, 55, 55, 32), True),
             ((64, 256, 27, 27, 160), False), ((16, 128, 30, 30, 96), False), ((64, 16, 33, 33, 32), False),
             ((32, 96, 56, 56, 96), True), ((16, 512, 112, 112, 96), True)]
    for (n, c, h, w, k), want in cases:
        rc, ki, prm = _ki(wc, wc.make_problem(n, c, h, w, k, ptrs=FAKE))
        assert rc == wc.OK and prm.form == wc.FORM_FUSED
        half = n * prm.fnbh * prm.fnbw * -(-k // 64)
        assert (ki.slot(2)["grid"][0] < half) == want, (n, c, h, w, k)


def test_getkernelinfo_fills_slots_and_param(wc):
    pb = wc.make_problem(32, 256, 56, 56, 256, ptrs=FAKE)
    rc, ki, prm = _ki(wc, pb)
    assert rc == wc.OK
    assert (prm.n, prm.c, prm.h, prm.w, prm.k, prm.r, prm.s, prm.u, prm.v, prm.p, prm.q) == (32, 256, 56, 56, 256, 3, 3, 1, 1, 1, 1)
    assert (prm.Oh, prm.Ow, prm.tiles_h, prm.tiles_w, prm.T, prm.NT) == (56, 56, 28, 28, 784, 25088)
    assert (prm.G32, prm.CB, prm.KB32) == (784, 8, 8)
    assert prm.pin == 256 and prm.transform_filter == 256 and prm.flags == 1
    for i in range(1, 5):
        s = ki.slot(i)
        assert s["kernel"] and all(s["grid"]) and all(s["block"])
    # this layer (K = 256 <= 320) takes the fused form: filter transform, fused kernel, two empty
    assert prm.form == wc.FORM_FUSED
    assert ki.slot(1)["grid"] == (prm.KB32 * 2 * prm.CB * 4, 1, 1) and ki.slot(1)["block"] == (64, 1, 1)
    assert ki.slot(2)["grid"] == (32 * prm.fnbh * prm.fnbw * 4, 1, 1) and ki.slot(2)["block"] == (512, 1, 1)
    assert ki.slot(3)["grid"] == (1, 1, 1) and ki.slot(4)["grid"] == (1, 1, 1)


def test_getkernelinfo_staged_slots(wc):
    """A staged layer (C = K = 512, N*T*K*C = 3.3e9): input transform, GEMM, empty."""
    pb = wc.make_problem(64, 512, 28, 28, 512, ptrs=FAKE)
    rc, ki, prm = _ki(wc, pb)
    assert rc == wc.OK and prm.form == wc.FORM_STAGED
    assert (prm.tiles_h, prm.tiles_w, prm.T, prm.NT, prm.G32, prm.CB, prm.KB32) == (14, 14, 196, 12544, 392, 16, 16)
    assert ki.slot(4)["grid"] == (1, 1, 1)
    cfg = wc.tile_configs()[prm.cfg]
    assert ki.slot(3)["block"][0] == cfg["threads"]
    # one-dimensional grid over (column block, k-block) pairs, remapped per XCD in the kernel
    assert ki.slot(3)["grid"] == (-(-prm.G32 // (cfg["bt"] // 32)) * -(-prm.KB32 // (cfg["bk"] // 32)), 1, 1)
    assert ki.slot(2)["grid"] == (-(-14 // prm.band), 16, 64)
    assert ki.slot(1)["grid"] == (prm.KB32 * 2 * prm.CB * 4, 1, 1) and ki.slot(1)["block"] == (64, 1, 1)


def test_weight_hash_is_fnv1a64():
    """The published FNV-1a 64 test vectors (offset basis, "a", "foobar")."""
    import importlib
    wc = importlib.import_module("2024pra_conv_amd")
    L = wc.load_library()
    h = ctypes.c_ulonglong()
    for data, want in ((b"", 0xcbf29ce484222325), (b"a", 0xaf63dc4c8601ec8c), (b"foobar", 0x85944171f73967e8)):
        buf = ctypes.create_string_buffer(data, len(data)) if data else None
        assert L.conv2d_weight_hash(buf, len(data), ctypes.byref(h)) == wc.OK and h.value == want
    assert L.conv2d_weight_hash(None, 4, ctypes.byref(h)) == wc.ERR_NULL


def test_weights_load_refuses_foreign_files_before_device_access(wc, tmp_path):
    """Header checks run on the host only (the problem's pointers here are not device memory,
    so any device access would fail with CONV2D_ERR_HIP instead)."""
    import struct
    L = wc.load_library()
    n, c, h, w, k = 2, 40, 10, 10, 24
    pb = wc.make_problem(n, c, h, w, k, ptrs=FAKE)
    _, ub, _ = wc.form_workspace_size(pb, wc.FORM_STAGED)  # the U of the problem's (F(2x2)) form

    def header(magic=b"WINOU16\0", version=1, method=wc.METHOD_WINOGRAD, kk=k, cc=c, r=3, s_=3, ubytes=ub, zero=0):
        return struct.pack("<8sII4IQQQQ", magic, version, method, kk, cc, r, s_, 0, ubytes, 0, zero)

    def load(data):
        f = tmp_path / "u.bin"
        f.write_bytes(data)
        return L.conv2d_weights_load(ctypes.byref(pb), str(f).encode(), None)

    assert L.conv2d_weights_load(ctypes.byref(pb), str(tmp_path / "missing.bin").encode(), None) == wc.ERR_IO
    assert load(header()[:40]) == wc.ERR_IO                           # truncated header
    assert load(header(magic=b"WINOU08\0")) == wc.ERR_MISMATCH
    assert load(header(version=2)) == wc.ERR_MISMATCH
    assert load(header(method=wc.METHOD_DIRECT)) == wc.ERR_MISMATCH
    assert load(header(method=2)) == wc.ERR_MISMATCH                  # an F(4x4) U
    assert load(header(kk=k + 1)) == wc.ERR_MISMATCH
    assert load(header(ubytes=ub - 1024)) == wc.ERR_MISMATCH
    assert load(header(zero=1)) == wc.ERR_MISMATCH
    assert load(header() + b"\0" * (ub - 1)) == wc.ERR_IO               # truncated payload
    assert L.conv2d_weights_load(None, b"x", None) == wc.ERR_NULL
    assert wc.load_library().conv2d_weights_save(None, b"x", None) == wc.ERR_NULL


def test_form_choice_and_fused_slots(wc):
    """Layers with K <= 384, N*T*K*C < 8e8, or N*T*K*C < 1.2e9 with C >= K take the fused
    form (slot 2 = the fused kernel, slots 3 and 4 empty), odd widths included (but not a
    one-pixel width); K >= 416 beyond that (28x28 C = K = 512 at N = 64, 56x56 C = 512
    K = 385, C < K = 512 at 8.2e8) and the direct path stay staged."""
    L = wc.load_library()
    f = ctypes.c_int()
    cases = [((32, 256, 56, 56, 256), {}, wc.FORM_FUSED), ((8, 256, 14, 14, 256), {}, wc.FORM_FUSED),
             ((64, 512, 28, 28, 512), {}, wc.FORM_STAGED), ((16, 512, 112, 112, 96), {}, wc.FORM_FUSED),
             ((32, 512, 56, 56, 384), {}, wc.FORM_FUSED), ((32, 512, 56, 56, 385), {}, wc.FORM_STAGED),
             ((32, 96, 55, 55, 96), {}, wc.FORM_FUSED), ((1, 64, 55, 55, 64), {}, wc.FORM_FUSED),
             ((4, 512, 13, 13, 96), {}, wc.FORM_FUSED), ((32, 256, 55, 55, 256), {}, wc.FORM_FUSED),
             ((16, 512, 28, 28, 512), {}, wc.FORM_FUSED), ((8, 512, 28, 28, 512), {}, wc.FORM_FUSED),
             ((32, 512, 28, 28, 512), {}, wc.FORM_STAGED), ((32, 256, 28, 28, 512), {}, wc.FORM_STAGED),
             ((32, 384, 56, 56, 384), {}, wc.FORM_FUSED), ((64, 256, 56, 56, 257), {}, wc.FORM_FUSED),
             ((2, 16, 9, 9, 16), dict(u=2, v=2), wc.FORM_STAGED),
             # a one-pixel-wide image: its 4-byte pair loads would start before the image
             ((4, 16, 9, 1, 16), {}, wc.FORM_STAGED), ((4, 16, 9, 3, 16), {}, wc.FORM_FUSED)]
    for (n, c, h, w, k), extra, want in cases:
        pb = wc.make_problem(n, c, h, w, k, **extra, ptrs=FAKE)
        assert L.conv2d_form(ctypes.byref(pb), ctypes.byref(f)) == wc.OK and f.value == want, (n, c, h, k)
        rc, ki, prm = _ki(wc, pb)
        assert rc == wc.OK and prm.form == want
        if want == wc.FORM_FUSED:
            # a last block of 1..32 channels runs as wide items (32 channels x two 64-tile blocks
            # side by side) with pair or quad staging and a block other than 2 x 32,
            # when the list schedule on the 256 compute units assumed without a GPU ends sooner
            half = n * prm.fnbh * prm.fnbw * -(-k // 64)
            wide = n * prm.fnbh * (prm.fnbw * (k // 64) + (prm.fnbw + 1) // 2)
            full = n * prm.fnbh * prm.fnbw * (k // 64)
            use_wide = (0 < k % 64 <= 32 and prm.fnu in (1, 2) and (prm.fbh, prm.fbw) != (2, 32)
                        and _list_schedule(full, 1.0, wide - full, 1.5) < _list_schedule(full, 1.0, half - full, 1.0))
            assert ki.slot(2)["block"] == (512, 1, 1) and ki.slot(2)["grid"][0] == (wide if use_wide else half)
            assert ki.slot(3)["grid"] == (1, 1, 1) and ki.slot(4)["grid"] == (1, 1, 1)
            assert ki.slot(2)["lds"] >= 131072
    assert L.conv2d_form(None, ctypes.byref(f)) == wc.ERR_NULL


def test_set_epilogue_fills_the_blob(wc):
    L = wc.load_library()
    pb = wc.make_problem(2, 16, 8, 8, 24, ptrs=FAKE)
    ki = wc.KernelInfo()
    buf = ctypes.create_string_buffer(ctypes.sizeof(wc.ParamType))
    assert L.getkernelInfo(ctypes.byref(pb), ctypes.byref(ki), buf) == wc.OK
    prm = wc.ParamType.from_buffer_copy(buf)
    assert prm.act == wc.ACT_NONE and not prm.bias  # no epilogue unless asked for
    assert L.conv2d_set_epilogue(buf, 4096, wc.ACT_RELU) == wc.OK
    prm = wc.ParamType.from_buffer_copy(buf)
    assert prm.act == wc.ACT_RELU and prm.bias == 4096 and prm.k == 24
    assert L.conv2d_set_epilogue(buf, None, 7) == wc.ERR_INVALID
    assert L.conv2d_set_epilogue(None, None, wc.ACT_NONE) == wc.ERR_NULL
    assert L.conv2d_forward_epilogue(ctypes.byref(pb), -1, None, 3, None) == wc.ERR_INVALID


@pytest.mark.parametrize("change,code", [
    ({"n": 0}, "ERR_INVALID"), ({"c": 40000}, "ERR_OVERFLOW"), ({"h": 1, "p": 0}, "ERR_INVALID"),
    ({"r": 0}, "ERR_INVALID"), ({"u": 0}, "ERR_INVALID"), ({"h": 2, "r": 5, "p": 1}, "ERR_INVALID"),
    # direct path: reduction index beyond 2^30, packed operand beyond 64 GiB
    ({"c": 30000, "r": 200, "s": 200, "p": 100, "q": 100}, "ERR_OVERFLOW"),
    ({"n": 8000, "c": 1024, "h": 256, "w": 256, "p": 3, "q": 3}, "ERR_OVERFLOW"),
])
def test_getkernelinfo_rejects(wc, change, code):
    args = dict(n=1, c=16, h=8, w=8, k=16, r=3, s=3, u=1, v=1, p=1, q=1)
    args.update(change)
    pb = wc.make_problem(**args, ptrs=FAKE)
    rc, _, _ = _ki(wc, pb)
    assert rc == getattr(wc, code)


@pytest.mark.parametrize("c", [40, 13])
@pytest.mark.parametrize("geo", [
    dict(r=5, s=5, p=2, q=2), dict(u=2, v=2), dict(p=3, q=3), dict(r=1, s=1, p=0, q=0),
    dict(r=1, s=1, u=2, v=2, p=0, q=0), dict(r=1, s=1, p=1, q=1), dict(r=7, s=7, u=2, v=2, p=3, q=3), dict(r=3, s=1, p=1, q=0), dict(r=2, s=2, u=2, v=2, p=0, q=0),
])
def test_general_geometry_runs_on_the_direct_path(wc, geo, c):
    """Geometry the Winograd path does not take (the reference ignored r, s, u, v, p, q and
    computed garbage) is accepted and set up as an implicit-GEMM direct convolution."""
    L = wc.load_library()
    args = dict(n=2, c=c, h=17, w=13, k=24, r=3, s=3, u=1, v=1, p=1, q=1)
    args.update(geo)
    pb = wc.make_problem(**args, ptrs=FAKE)
    m = ctypes.c_int()
    assert L.conv2d_method(ctypes.byref(pb), ctypes.byref(m)) == wc.OK and m.value == wc.METHOD_DIRECT
    rc, ki, prm = _ki(wc, pb)
    assert rc == wc.OK
    oh = (args["h"] - args["r"] + 2 * args["p"]) // args["u"] + 1
    ow = (args["w"] - args["s"] + 2 * args["q"]) // args["v"] + 1
    cb = -(-args["c"] // 32)
    P = args["n"] * oh * ow
    nsteps = -(-(-(-(args["r"] * args["s"] * args["c"]) // 32)) // 8) * 8
    assert (prm.Oh, prm.Ow, prm.T, prm.NT, prm.G32, prm.CB, prm.nsteps) == (oh, ow, oh * ow, P, -(-P // 32), cb, nsteps)
    assert prm.KB32 == -(-args["k"] // 32)
    vb, ub, mb = wc.workspace_size(pb)
    gather = args["r"] == 1 and args["s"] == 1
    # a 1x1 filter: the GEMM gathers the input itself, so there is no packed input and slot 2
    # (the packing pass otherwise) is the empty kernel of slot 4 (larger filters gather only
    # from channel-contiguous input)
    assert vb == (0 if gather else nsteps * 2 * prm.G32 * 1024)
    assert ub == nsteps * 2 * prm.KB32 * 1024 and mb == 0
    assert all(ki.slot(i)["kernel"] for i in range(1, 5))
    assert ki.slot(1)["grid"][0] * 256 >= nsteps * 2 * prm.KB32 * 64
    if gather:
        assert ki.slot(2)["kernel"] == ki.slot(4)["kernel"] and ki.slot(2)["grid"] == (1, 1, 1)
    elif args["c"] % 8:
        # few channels: one wave per 32-pixel block, walking every 16-value block
        assert ki.slot(2)["grid"] == (-(-prm.G32 // 4), 1, 1) and ki.slot(2)["block"] == (256, 1, 1)
    else:
        assert ki.slot(2)["grid"][0] * 256 >= nsteps * 2 * prm.G32 * 64
        assert ki.slot(2)["kernel"] != ki.slot(4)["kernel"]
    ncfg = L.conv2d_direct_num_tile_configs()
    assert 0 <= prm.cfg < ncfg
    bk, bt, th = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert L.conv2d_direct_tile_config(prm.cfg, ctypes.byref(bk), ctypes.byref(bt), ctypes.byref(th)) == wc.OK
    assert ki.slot(3)["grid"] == (-(-prm.G32 // (bt.value // 32)) * -(-prm.KB32 // (bk.value // 32)), 1, 1)
    assert ki.slot(3)["block"][0] == th.value
    # the Winograd-only forms refuse it
    a, b = ctypes.c_size_t(), ctypes.c_size_t()
    assert L.conv2d_fp8_workspace_size(ctypes.byref(pb), ctypes.byref(a), ctypes.byref(b)) == wc.ERR_UNSUPPORTED


def test_winograd_geometry_reports_winograd(wc):
    L = wc.load_library()
    m = ctypes.c_int()
    for p in (0, 1, 2):
        pb = wc.make_problem(1, 16, 9, 9, 16, p=p, q=p)
        assert L.conv2d_method(ctypes.byref(pb), ctypes.byref(m)) == wc.OK and m.value == wc.METHOD_WINOGRAD


@pytest.mark.parametrize("shape,in_addr,fnu,fsh,fpc", [
    ((32, 256, 56, 56, 256, 1, 1), 256, 1, 3, 40),  # width 4k, 8-byte base: quads, shift 3 (4x16)
    ((1, 32, 16, 16, 64, 1, 0), 256, 1, 0, 20),     # q = 0: quads, shift 0
    ((1, 32, 16, 16, 64, 1, 2), 256, 1, 2, 20),     # q = 2: quads, shift 2
    ((2, 40, 14, 22, 70, 1, 1), 256, 2, 1, 20),     # width 2 mod 4: pixel pairs, shift q & 1
    ((32, 256, 56, 56, 256, 1, 1), 260, 2, 1, 36),  # base 4 mod 8: pixel pairs
    ((32, 256, 56, 56, 256, 1, 1), 258, 2, 1, 36),  # base 2 mod 4: 2-byte pixels
    ((1, 33, 70, 8, 16, 1, 1), 256, 2, 1, 12),      # 16x4 block, shift 3: 544 quads > 512 -> pairs
    ((1, 33, 70, 4, 16, 1, 2), 256, 1, 2, 12),      # 16x4 block, shift 2: 408 quads
])
def test_fused_staging_width(wc, shape, in_addr, fnu, fsh, fpc):
    """The fused form stages the input patch with 8-byte pixel-quad loads (one unit per
    thread, fnu = 1) when the width is a multiple of 4, the input base 8-byte aligned and the
    quad patch fits 512 units; the patch then starts (4 - q % 4) % 4 columns left of 2*tw0 - q
    and is a multiple of 4 columns wide. Otherwise 4-byte pairs (or 2-byte pixels), two units
    per thread, shift q & 1, 2*fbw + 4 columns."""
    n, c, h, w, k, p, q = shape
    pb = wc.make_problem(n, c, h, w, k, p=p, q=q, ptrs=dict(FAKE, in_=in_addr))
    rc, ki, prm = _ki(wc, pb)
    assert rc == wc.OK and prm.form == wc.FORM_FUSED
    assert (prm.fnu, prm.fsh, prm.fpc) == (fnu, fsh, fpc)
    assert prm.fpr == 2 * prm.fbh + 2
    assert ki.slot(2)["lds"] >= 2 * 4 * prm.fpr * prm.fpc * 16


def _ki_io(wc, pb, io):
    L = wc.load_library()
    ki = wc.KernelInfo()
    buf = ctypes.create_string_buffer(ctypes.sizeof(wc.ParamType))
    rc = L.getkernelInfo_io(ctypes.byref(pb), io, ctypes.byref(ki), buf)
    return rc, ki, wc.ParamType.from_buffer_copy(buf)


@pytest.mark.parametrize("shape,fpc,fnu", [
    ((1, 32, 16, 16, 64, 1, 1), 18, 3),      # 8x8 block: 4 x 18 x 18 = 1296 single-pixel units
    ((32, 256, 56, 56, 256, 1, 1), 34, 3),   # 4x16 block (28x28 tiles): 4 x 10 x 34 = 1360
    ((1, 32, 70, 6, 16, 1, 0), 10, 3),       # 16x4 block: 4 x 34 x 10 = 1360
    ((1, 8, 4, 130, 32, 1, 1), 66, 4),       # 2x32 block (2x65 tiles): 4 x 6 x 66 = 1584
])
def test_nhwc_input_staging_geometry(wc, shape, fpc, fnu):
    """Channel-contiguous input on the fused form: one 16-byte unit (8 channels of a pixel)
    per load, no column shift, 2*fbw + 2 patch columns, up to 4 units per thread; io is
    carried in the blob, and output-only bits leave the NCHW staging as it is."""
    n, c, h, w, k, p, q = shape
    pb = wc.make_problem(n, c, h, w, k, p=p, q=q, ptrs=FAKE)
    io = wc.IO_IN_NHWC | wc.IO_OUT_NHWC | wc.IO_OUT_BF16
    rc, ki, prm = _ki_io(wc, pb, io)
    assert rc == wc.OK and prm.form == wc.FORM_FUSED and prm.io == io
    assert (prm.fsh, prm.fpc, prm.fnu) == (0, fpc, fnu)
    assert prm.fpc == 2 * prm.fbw + 2 and prm.fpr == 2 * prm.fbh + 2
    _, _, nchw = _ki(wc, pb)
    rc, _, outonly = _ki_io(wc, pb, wc.IO_OUT_NHWC | wc.IO_OUT_BF16)
    assert rc == wc.OK and nchw.io == 0
    assert (outonly.fsh, outonly.fpc, outonly.fnu, outonly.form) == (nchw.fsh, nchw.fpc, nchw.fnu, nchw.form)


def test_nhwc_input_without_whole_channel_groups_is_staged(wc):
    """The fused form's NHWC loads need C % 8 == 0 and a 16-byte aligned input; otherwise the
    layer runs staged (the input transform reads any C and alignment), and bits beyond the
    three io flags are refused."""
    for c, addr in ((36, 256), (256, 264)):
        pb = wc.make_problem(32, c, 56, 56, 256, ptrs=dict(FAKE, in_=addr))
        rc, ki, prm = _ki_io(wc, pb, wc.IO_IN_NHWC)
        assert rc == wc.OK and prm.form == wc.FORM_STAGED and prm.io == wc.IO_IN_NHWC, c
        assert ki.slot(3)["grid"][0] > 1
        rc, _, prm = _ki_io(wc, pb, wc.IO_NCHW)
        assert rc == wc.OK and prm.form == wc.FORM_FUSED
    pb = wc.make_problem(32, 256, 56, 56, 256, ptrs=FAKE)
    for bad in (8, 16, -1):
        assert _ki_io(wc, pb, bad)[0] == wc.ERR_INVALID
    L = wc.load_library()
    assert L.conv2d_forward_io(ctypes.byref(pb), -1, 8, None) == wc.ERR_INVALID


@pytest.mark.parametrize("w,band", [(224, 1), (512, 1), (629, 1), (631, 0), (1500, 0)])
def test_wide_images_stage_one_tile_row(wc, w, band):
    """The input transform stages whole input rows: one tile row (4 rows x 32 channels) of a
    w-wide image needs 256*(w + 9) bytes of LDS, so up to w = 629 (160 KiB, one workgroup per
    compute unit beyond 64 KiB) both forms apply; wider images stay on the Winograd path in the
    fused form only (its patches are per tile block), and the staged-only entry points refuse
    them before touching the device."""
    L = wc.load_library()
    m, f = ctypes.c_int(), ctypes.c_int()
    pb = wc.make_problem(1, 64, 6, w, 64, ptrs=FAKE)
    assert L.conv2d_method(ctypes.byref(pb), ctypes.byref(m)) == wc.OK and m.value == wc.METHOD_WINOGRAD
    rc, ki, prm = _ki(wc, pb)
    assert rc == wc.OK and prm.band == band
    if band == 0:
        assert prm.form == wc.FORM_FUSED and L.conv2d_form(ctypes.byref(pb), ctypes.byref(f)) == wc.OK
        assert f.value == wc.FORM_FUSED
        assert L.conv2d_input_transform(ctypes.byref(pb), None) == wc.ERR_UNSUPPORTED
        assert L.conv2d_gemm(ctypes.byref(pb), 0, None) == wc.ERR_UNSUPPORTED
        assert L.conv2d_forward(ctypes.byref(pb), 0, None) == wc.ERR_UNSUPPORTED
        vb, ub = ctypes.c_size_t(), ctypes.c_size_t()
        assert L.conv2d_fp8_workspace_size(ctypes.byref(pb), ctypes.byref(vb), ctypes.byref(ub)) == wc.ERR_UNSUPPORTED


def test_getkernelinfo_null_pointers(wc):
    L = wc.load_library()
    pb = wc.make_problem(1, 16, 8, 8, 16)  # no device pointers
    rc, _, _ = _ki(wc, pb)
    assert rc == wc.ERR_NULL
    assert L.getkernelInfo(None, None, None) == wc.ERR_NULL
    assert L.conv2d_launch(None, None, None) == wc.ERR_NULL
    ki = wc.KernelInfo()  # all-zero slots: refused before any launch
    assert L.conv2d_launch(ctypes.byref(ki), ctypes.c_void_p(1), None) == wc.ERR_NULL


def test_workspace_sizes_and_legacy_sizing(wc):
    # legacy harness sizes (main.cpp:36-38): V 4*h*h*n*c, U 16*k*c halves
    for (n, c, h, k) in [(32, 256, 56, 256), (1, 64, 56, 64), (64, 512, 28, 512), (32, 96, 55, 96),
                         (1, 3, 13, 32), (16, 128, 64, 27), (2, 40, 10, 20)]:
        pb = wc.make_problem(n, c, h, h, k)
        vb, ub, mb = wc.form_workspace_size(pb, wc.FORM_STAGED)
        tiles = (h + 1) // 2
        g32 = -(-(n * tiles * tiles) // 32)
        cb, kb = -(-c // 32), -(-k // 32)
        assert vb == 16 * 2 * cb * g32 * 512 * 2 and ub == 16 * 2 * cb * kb * 512 * 2 and mb == 0
        assert wc.form_workspace_size(pb, wc.FORM_FUSED) == (vb, ub, 0)
        # the F(4x4) form: 36 positions over 4x4 tiles; M as FP16 with rows and columns padded to 32
        t4 = (h + 3) // 4
        g4 = -(-(n * t4 * t4) // 32)
        v4, u4, m4 = wc.form_workspace_size(pb, wc.FORM_F43)
        assert (v4, u4, m4) == (36 * 2 * cb * g4 * 1024, 36 * 2 * cb * kb * 1024, 36 * kb * 32 * g4 * 32 * 2)
        # conv2d_workspace_size covers every form
        assert wc.workspace_size(pb) == (max(vb, v4), max(ub, u4), m4)
        legacy_ok = vb <= 4 * h * h * n * c * 2 and ub <= 16 * k * c * 2
        predicted = h % 2 == 0 and c % 32 == 0 and k % 32 == 0 and (n * tiles * tiles) % 32 == 0
        assert legacy_ok == predicted, (n, c, h, k)


def test_tile_configs_and_selection(wc):
    cfgs = wc.tile_configs()
    assert len(cfgs) >= 8
    L = wc.load_library()
    assert L.conv2d_tile_config(len(cfgs), None, None, None) == wc.ERR_INVALID
    for c in cfgs:
        assert c["bk"] % 16 == 0 and c["bt"] % 16 == 0 and c["threads"] % 64 == 0 and c["threads"] <= 1024


@pytest.mark.parametrize("shape,want", [((32, 512, 28, 28, 512), 9), ((8, 512, 56, 56, 512), 9),
                                        ((32, 256, 56, 56, 256), 9), ((16, 512, 28, 28, 512), 11),
                                        ((16, 128, 56, 56, 128), 11), ((4, 64, 28, 28, 128), 10),
                                        ((1, 64, 56, 56, 64), 10)])
def test_tile_heuristic(wc, shape, want):
    """Untuned shapes: 32x64 (configuration 10) below one 64x128 workgroup per two compute
    units; the 128x128 LDS-DMA tile (9) for C >= 256 with at least one of its workgroups per
    two compute units (196 workgroups for 32x512x28x28x512); else 64x128 (11)."""
    L = wc.load_library()
    L.conv2d_tuned_clear()
    pb = wc.make_problem(*shape, ptrs=FAKE)
    c = ctypes.c_int()
    assert L.conv2d_select_tile_config(ctypes.byref(pb), ctypes.byref(c)) == wc.OK and c.value == want


def test_tuned_table_roundtrip(wc, tmp_path):
    L = wc.load_library()
    L.conv2d_tuned_clear()
    pb = wc.make_problem(4, 64, 28, 28, 128, ptrs=FAKE)
    c = ctypes.c_int()
    assert L.conv2d_select_tile_config(ctypes.byref(pb), ctypes.byref(c)) == wc.OK
    default = c.value
    other = (default + 1) % L.conv2d_num_tile_configs()
    assert L.conv2d_tuned_set(ctypes.byref(pb), other) == wc.OK
    assert L.conv2d_select_tile_config(ctypes.byref(pb), ctypes.byref(c)) == wc.OK and c.value == other
    rc, _, prm = _ki(wc, pb)
    assert rc == wc.OK and prm.cfg == other
    path = str(tmp_path / "tuned.txt").encode()
    assert L.conv2d_tuned_save(path) == wc.OK
    L.conv2d_tuned_clear()
    assert L.conv2d_tuned_count() == 0
    assert L.conv2d_tuned_load(path) == wc.OK and L.conv2d_tuned_count() == 1
    assert L.conv2d_select_tile_config(ctypes.byref(pb), ctypes.byref(c)) == wc.OK and c.value == other
    assert L.conv2d_tuned_set(ctypes.byref(pb), 999) == wc.ERR_INVALID
    bad = tmp_path / "bad.txt"
    bad.write_text("1 2 3\n")
    assert L.conv2d_tuned_load(str(bad).encode()) == wc.ERR_INVALID
    L.conv2d_tuned_clear()


def test_tuned_table_records_the_form(wc, tmp_path):
    """A tuned entry carries the form: getkernelInfo fills the fused slots for a fused entry and
    the staged ones for a staged entry, whatever the hand rule would pick; the form survives a
    save/load round trip, and a file without the form column loads as staged."""
    L = wc.load_library()
    L.conv2d_tuned_clear()
    pb = wc.make_problem(32, 256, 56, 56, 256, ptrs=FAKE)
    f = ctypes.c_int()
    assert L.conv2d_form(ctypes.byref(pb), ctypes.byref(f)) == wc.OK and f.value == wc.FORM_FUSED  # K <= 384
    assert L.conv2d_tuned_set_form(ctypes.byref(pb), 9, wc.FORM_STAGED) == wc.OK
    assert L.conv2d_form(ctypes.byref(pb), ctypes.byref(f)) == wc.OK and f.value == wc.FORM_STAGED
    rc, ki, prm = _ki(wc, pb)
    assert rc == wc.OK and prm.form == wc.FORM_STAGED and prm.cfg == 9
    assert L.conv2d_tuned_set_form(ctypes.byref(pb), 9, wc.FORM_FUSED) == wc.OK
    rc, ki, prm = _ki(wc, pb)
    assert rc == wc.OK and prm.form == wc.FORM_FUSED
    path = str(tmp_path / "tuned.txt").encode()
    assert L.conv2d_tuned_save(path) == wc.OK
    assert (tmp_path / "tuned.txt").read_text().splitlines()[1].split() == ["32", "256", "56", "56", "256", "1", "1", "9", "1"]
    L.conv2d_tuned_clear()
    assert L.conv2d_tuned_load(path) == wc.OK
    assert L.conv2d_form(ctypes.byref(pb), ctypes.byref(f)) == wc.OK and f.value == wc.FORM_FUSED
    old = tmp_path / "old.txt"
    old.write_text("32 256 56 56 256 1 1 9\n")
    L.conv2d_tuned_clear()
    assert L.conv2d_tuned_load(str(old).encode()) == wc.OK
    assert L.conv2d_form(ctypes.byref(pb), ctypes.byref(f)) == wc.OK and f.value == wc.FORM_STAGED
    assert L.conv2d_tuned_set_form(ctypes.byref(pb), 9, 7) == wc.ERR_INVALID
    bad = tmp_path / "bad.txt"
    bad.write_text("32 256 56 56 256 1 1 9 5\n")
    assert L.conv2d_tuned_load(str(bad).encode()) == wc.ERR_INVALID
    L.conv2d_tuned_clear()


def test_missing_library_fails_loudly(wc, tmp_path):
    with pytest.raises(RuntimeError):
        wc.load_library(str(tmp_path / "nope.so"))


def test_f43_form_slots_from_the_tuned_table(wc, tmp_path):
    """A shape recorded as CONV2D_FORM_F43 runs its four slots (filter, input transform, the
    36-position GEMM, output transform) when mid_result is set; without mid_result the shape
    falls back to the heuristic form. The blob carries the 4x4-tile geometry."""
    L = wc.load_library()
    f = ctypes.c_int()
    ptrs = dict(FAKE, mid_result=256)
    pb = wc.make_problem(32, 256, 56, 56, 256, ptrs=ptrs)
    L.conv2d_tuned_clear()
    try:
        assert L.conv2d_tuned_set_form(ctypes.byref(pb), 5, wc.FORM_F43) == wc.OK
        assert L.conv2d_form(ctypes.byref(pb), ctypes.byref(f)) == wc.OK and f.value == wc.FORM_F43
        rc, ki, prm = _ki(wc, pb)
        assert rc == wc.OK and prm.form == wc.FORM_F43 and prm.cfg == 5
        assert (prm.tiles_h, prm.tiles_w, prm.T, prm.NT, prm.G32, prm.CB, prm.KB32) == (14, 14, 196, 6272, 196, 8, 8)
        # band: 4 tile rows need 18 x 66 x 64 B > 64 KiB of LDS, 2 rows 42,240 B
        assert prm.band == 2 and ki.slot(2)["lds"] == 32 * 10 * 66 * 2
        assert ki.slot(1)["grid"] == (8 * 16 * 512 // 256, 1, 1) and ki.slot(1)["block"] == (256, 1, 1)
        assert ki.slot(2)["grid"] == (7, 8, 32)
        bk, bt, threads = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        assert L.conv2d_f43_tile_config(5, ctypes.byref(bk), ctypes.byref(bt), ctypes.byref(threads)) == wc.OK
        # positions per GEMM workgroup: the most (a divisor of 36) keeping two workgroups per
        # compute unit (256 when no GPU is visible)
        base = -(-8 // (bk.value // 32)) * -(-196 // (bt.value // 32))
        npos = next(p for p in (36, 18, 12, 9, 6, 4, 3, 2, 1) if base * (36 // p) >= 512)
        assert prm.f43_npos == npos == 6
        assert ki.slot(3)["grid"] == (base * (36 // npos), 1, 1)
        assert ki.slot(3)["block"] == (threads.value, 1, 1)
        assert ki.slot(4)["grid"] == (-(-6272 // 256), 64, 1) and ki.slot(4)["block"] == (256, 1, 1)
        # the tuned file keeps the form
        path = str(tmp_path / "t.txt").encode()
        assert L.conv2d_tuned_save(path) == wc.OK and L.conv2d_tuned_clear() == wc.OK
        assert L.conv2d_tuned_load(path) == wc.OK and L.conv2d_tuned_count() == 1
        assert L.conv2d_form(ctypes.byref(pb), ctypes.byref(f)) == wc.OK and f.value == wc.FORM_F43
        # no mid_result: the heuristic form (fused for K <= 384)
        nomid = wc.make_problem(32, 256, 56, 56, 256, ptrs=FAKE)
        assert L.conv2d_form(ctypes.byref(nomid), ctypes.byref(f)) == wc.OK and f.value == wc.FORM_FUSED
        rc, _, prm = _ki(wc, nomid)
        assert rc == wc.OK and prm.form == wc.FORM_FUSED
        # an F(4x4) entry indexes the F(4x4) GEMM table
        n43 = L.conv2d_f43_num_tile_configs()
        assert L.conv2d_tuned_set_form(ctypes.byref(pb), n43, wc.FORM_F43) == wc.ERR_INVALID
        assert L.conv2d_tuned_set_form(ctypes.byref(pb), n43, wc.FORM_STAGED) == wc.OK
    finally:
        L.conv2d_tuned_clear()


def test_f43_entry_points_refuse(wc):
    L = wc.load_library()
    pb = wc.make_problem(2, 16, 12, 12, 16, ptrs=FAKE)  # no mid_result
    assert L.conv2d_f43_gemm(ctypes.byref(pb), -1, None) == wc.ERR_NULL
    assert L.conv2d_f43_output_transform(ctypes.byref(pb), None) == wc.ERR_NULL
    assert L.conv2d_f43_forward(ctypes.byref(pb), -1, None) == wc.ERR_NULL
    withmid = wc.make_problem(2, 16, 12, 12, 16, ptrs=dict(FAKE, mid_result=256))
    assert L.conv2d_f43_gemm(ctypes.byref(withmid), L.conv2d_f43_num_tile_configs(), None) == wc.ERR_INVALID
    direct = wc.make_problem(2, 16, 12, 12, 16, r=5, s=5, p=2, q=2, ptrs=dict(FAKE, mid_result=256))
    for fn in (L.conv2d_f43_prepare_weights, L.conv2d_f43_input_transform, L.conv2d_f43_output_transform):
        assert fn(ctypes.byref(direct), None) == wc.ERR_UNSUPPORTED
    assert wc.form_workspace_size(direct, wc.FORM_F43) is None
    assert L.conv2d_tuned_set_form(ctypes.byref(pb), 0, 3) == wc.ERR_INVALID
    assert L.conv2d_f43_tile_config(-1, None, None, None) == wc.ERR_INVALID


def test_null_input_workspace_only_where_it_is_zero_bytes(wc):
    """hipMalloc(0) gives NULL: a problem whose reported input workspace is 0 bytes (the direct
    path gathering its input: a 1x1 filter) may pass transform_pin = NULL; others get
    CONV2D_ERR_NULL."""
    no_v = dict(FAKE, transform_pin=None)
    cases = [(dict(n=2, c=16, h=9, w=9, k=8, r=1, s=1, p=0, q=0), wc.OK),          # 1x1
             (dict(n=2, c=13, h=9, w=9, k=8, r=1, s=1, u=2, v=2, p=0, q=0), wc.OK),  # 1x1 stride 2
             (dict(n=2, c=16, h=9, w=9, k=8, r=3, s=3, u=2, v=2), wc.ERR_NULL),    # packed im2col
             (dict(n=2, c=16, h=9, w=9, k=8), wc.ERR_NULL)]                        # Winograd
    for args, want in cases:
        pb = wc.make_problem(**args, ptrs=no_v)
        vb, _, _ = wc.workspace_size(pb)
        assert (vb == 0) == (want == wc.OK)
        rc, _, _ = _ki(wc, pb)
        assert rc == want, args
