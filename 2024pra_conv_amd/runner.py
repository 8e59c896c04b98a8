"""Drives the C-ABI library on PyTorch-allocated device buffers (tests and benchmarks).

PyTorch provides device memory and the stream handle only; every kernel is the library's.
"""
from __future__ import annotations

import ctypes

import torch

from . import (FORM_F43, FORM_STAGED, KernelInfo, ParamType, check, form_workspace_size, fp8_workspace_size,
               load_library, make_problem, workspace_size)


def _stream_handle(stream: torch.cuda.Stream | None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class WinogradConv:
    """One convolution problem with its own output and workspaces.

    x: (N, C, H, W) float16 on cuda; w: (K, C, R, S) float16 on cuda. 3x3 stride-1 problems
    with pad <= 2 run on the Winograd path, any other geometry on the direct path (the
    library decides; `method()` reports which).
    """

    def __init__(self, x: torch.Tensor, w: torch.Tensor, pad=(1, 1), out: torch.Tensor | None = None,
                 workspace: tuple[torch.Tensor, torch.Tensor] | None = None, stride=(1, 1), io: int = 0):
        """io: CONV2D_IO_* bits; with IO_IN_NHWC x is (N, H, W, C), with IO_OUT_NHWC the output is
        (N, Oh, Ow, K), with IO_OUT_BF16 it is bfloat16."""
        assert x.dtype == torch.float16 and w.dtype == torch.float16
        assert x.is_cuda and w.is_cuda and x.is_contiguous() and w.is_contiguous()
        self.L = load_library()
        self.io = io
        if io & 1:
            n, h, wd, c = x.shape
        else:
            n, c, h, wd = x.shape
        k, c2, r, s = w.shape
        assert c == c2
        self.x, self.w = x, w
        self.pad = pad
        self.shape = (n, c, h, wd, k, r, s, stride[0], stride[1], pad[0], pad[1])
        self.oh = (h - r + 2 * pad[0]) // stride[0] + 1
        self.ow = (wd - s + 2 * pad[1]) // stride[1] + 1
        probe = make_problem(*self.shape)
        vb, ub, mb = workspace_size(probe)  # enough for every form
        # the F(2x2) (or direct-path) operand sizes, and the F(4x4) form's where it applies
        self.v_bytes, self.u_bytes, _ = form_workspace_size(probe, FORM_STAGED)
        f43 = form_workspace_size(probe, FORM_F43)
        self.v43_bytes, self.u43_bytes, self.m43_bytes = f43 if f43 else (0, 0, 0)
        dev = x.device
        oshape = (n, self.oh, self.ow, k) if io & 2 else (n, k, self.oh, self.ow)
        odt = torch.bfloat16 if io & 4 else torch.float16
        self.out = out if out is not None else torch.empty(oshape, dtype=odt, device=dev)
        if workspace is not None:
            self.V, self.U = workspace[0], workspace[1]
            self.M = workspace[2] if len(workspace) > 2 else None
        else:
            self.V = torch.empty(max(vb, 2), dtype=torch.uint8, device=dev)
            self.U = torch.empty(max(ub, 2), dtype=torch.uint8, device=dev)
            self.M = torch.empty(mb, dtype=torch.uint8, device=dev) if mb else None
        self.pb = make_problem(*self.shape, ptrs={
            "in_": x.data_ptr(), "weight": w.data_ptr(), "out": self.out.data_ptr(),
            "transform_pin": self.V.data_ptr(), "transform_filter": self.U.data_ptr(),
            "mid_result": self.M.data_ptr() if self.M is not None else None})
        self._ki = None
        self._param = None

    # --- reference protocol: getParamsize / getkernelInfo / four slots -----------------
    def kernel_info(self):
        if self._ki is None:
            size = ctypes.c_int()
            check(self.L.getParamsize(ctypes.byref(self.pb), ctypes.byref(size)), "getParamsize")
            assert size.value == ctypes.sizeof(ParamType)
            self._param = ctypes.create_string_buffer(size.value)
            ki = KernelInfo()
            check(self.L.getkernelInfo_io(ctypes.byref(self.pb), self.io, ctypes.byref(ki), self._param),
                  "getkernelInfo_io")
            self._ki = ki
        return self._ki, self._param

    def param(self) -> ParamType:
        _, buf = self.kernel_info()
        return ParamType.from_buffer_copy(buf)

    def launch_slots(self, stream=None) -> None:
        ki, param = self.kernel_info()
        check(self.L.conv2d_launch(ctypes.byref(ki), param, _stream_handle(stream)), "conv2d_launch")

    # --- stage entry points -------------------------------------------------------------
    def prepare_weights(self, stream=None) -> None:
        check(self.L.conv2d_prepare_weights(ctypes.byref(self.pb), _stream_handle(stream)), "conv2d_prepare_weights")

    def input_transform(self, stream=None) -> None:
        check(self.L.conv2d_input_transform(ctypes.byref(self.pb), _stream_handle(stream)), "conv2d_input_transform")

    def gemm(self, cfg: int = -1, stream=None) -> None:
        check(self.L.conv2d_gemm(ctypes.byref(self.pb), cfg, _stream_handle(stream)), "conv2d_gemm")

    def forward(self, cfg: int = -1, stream=None) -> None:
        if self.io:
            check(self.L.conv2d_forward_io(ctypes.byref(self.pb), cfg, self.io, _stream_handle(stream)),
                  "conv2d_forward_io")
            return
        check(self.L.conv2d_forward(ctypes.byref(self.pb), cfg, _stream_handle(stream)), "conv2d_forward")

    def forward_epilogue(self, bias: torch.Tensor | None, act: int, cfg: int = -1, stream=None) -> None:
        """S2 + S3 with out = act(conv + bias[k]) (bias: K float16 on the device, or None)."""
        ptr = None if bias is None else bias.data_ptr()
        check(self.L.conv2d_forward_epilogue(ctypes.byref(self.pb), cfg, ptr, act, _stream_handle(stream)),
              "conv2d_forward_epilogue")

    def set_slot_epilogue(self, bias: torch.Tensor | None, act: int) -> None:
        """Makes the four slots (launch_slots) apply the epilogue."""
        _, param = self.kernel_info()
        ptr = None if bias is None else bias.data_ptr()
        check(self.L.conv2d_set_epilogue(param, ptr, act), "conv2d_set_epilogue")

    def save_weights(self, path: str, stream=None) -> None:
        """Writes the prepared U to a file keyed by the weights' hash (conv2d_weights_save)."""
        check(self.L.conv2d_weights_save(ctypes.byref(self.pb), path.encode(), _stream_handle(stream)),
              "conv2d_weights_save")

    def load_weights(self, path: str, stream=None) -> int:
        """Loads U from a file; returns the library's code (OK, ERR_MISMATCH, ERR_IO, ...)."""
        return self.L.conv2d_weights_load(ctypes.byref(self.pb), path.encode(), _stream_handle(stream))

    def form(self) -> int:
        """FORM_STAGED or FORM_FUSED: the form the four slots and forward() use."""
        f = ctypes.c_int()
        check(self.L.conv2d_form(ctypes.byref(self.pb), ctypes.byref(f)), "conv2d_form")
        return f.value

    def fused_gemm(self, stream=None) -> None:
        """Fused form: input transform inside the GEMM (needs prepare_weights; V unused)."""
        check(self.L.conv2d_fused_gemm(ctypes.byref(self.pb), _stream_handle(stream)), "conv2d_fused_gemm")

    def autotune(self, reps: int = 5, stream=None, forms: bool = False, f43: bool = False):
        """Runs the library's autotuner; returns (best_cfg, per-config best times in us, -1 = gated out)
        and, with forms=True, also (input-transform us, fused-kernel us or -1), with f43=True also
        the F(4x4) form's us (input transform + best GEMM + output transform) or -1."""
        n = self.L.conv2d_num_tile_configs()
        times = (ctypes.c_float * (n + 3))()
        best = ctypes.c_int(-1)
        check(self.L.conv2d_autotune(ctypes.byref(self.pb), reps, _stream_handle(stream), ctypes.byref(best),
                                     times, n + 3), "conv2d_autotune")
        t = list(times)
        if f43:
            return best.value, t[:n], t[n], t[n + 1], t[n + 2]
        return (best.value, t[:n], t[n], t[n + 1]) if forms else (best.value, t[:n])

    def selected_config(self) -> int:
        c = ctypes.c_int()
        check(self.L.conv2d_select_tile_config(ctypes.byref(self.pb), ctypes.byref(c)), "conv2d_select_tile_config")
        return c.value

    def method(self) -> int:
        m = ctypes.c_int()
        check(self.L.conv2d_method(ctypes.byref(self.pb), ctypes.byref(m)), "conv2d_method")
        return m.value

    def direct_f64(self, stream=None) -> torch.Tensor:
        """FP64 direct convolution, N,K,Oh,Ow (whatever the io layouts)."""
        n, k = self.shape[0], self.shape[4]
        ref = torch.empty((n, k, self.oh, self.ow), dtype=torch.float64, device=self.x.device)
        check(self.L.conv2d_direct_f64_io(ctypes.byref(self.pb), self.io, ctypes.c_void_p(ref.data_ptr()),
                                          _stream_handle(stream)), "conv2d_direct_f64_io")
        return ref

    def out_nchw(self) -> torch.Tensor:
        """The output as N,K,Oh,Ow (a permuted copy for NHWC output)."""
        return self.out.permute(0, 3, 1, 2).contiguous() if self.io & 2 else self.out

    # --- F(4x4,3x3) form stages (CONV2D_FORM_F43) -----------------------------------------
    def f43_prepare_weights(self, stream=None) -> None:
        check(self.L.conv2d_f43_prepare_weights(ctypes.byref(self.pb), _stream_handle(stream)),
              "conv2d_f43_prepare_weights")

    def f43_input_transform(self, stream=None) -> None:
        check(self.L.conv2d_f43_input_transform(ctypes.byref(self.pb), _stream_handle(stream)),
              "conv2d_f43_input_transform")

    def f43_gemm(self, cfg: int = -1, stream=None) -> None:
        check(self.L.conv2d_f43_gemm(ctypes.byref(self.pb), cfg, _stream_handle(stream)), "conv2d_f43_gemm")

    def f43_output_transform(self, stream=None) -> None:
        check(self.L.conv2d_f43_output_transform(ctypes.byref(self.pb), _stream_handle(stream)),
              "conv2d_f43_output_transform")

    def f43_forward(self, cfg: int = -1, stream=None) -> None:
        check(self.L.conv2d_f43_forward(ctypes.byref(self.pb), cfg, _stream_handle(stream)), "conv2d_f43_forward")

    def f43_selected_config(self) -> int:
        c = ctypes.c_int()
        check(self.L.conv2d_f43_select_tile_config(ctypes.byref(self.pb), ctypes.byref(c)),
              "conv2d_f43_select_tile_config")
        return c.value

    def set_form(self, form: int, cfg: int) -> None:
        """Records (form, cfg) for this shape in the library's tuned table."""
        check(self.L.conv2d_tuned_set_form(ctypes.byref(self.pb), cfg, form), "conv2d_tuned_set_form")

    # --- raw workspace views --------------------------------------------------------------
    def u_halves(self) -> torch.Tensor:
        return self.U[: self.u_bytes].view(torch.float16)

    def v_halves(self) -> torch.Tensor:
        return self.V[: self.v_bytes].view(torch.float16)

    def u43_halves(self) -> torch.Tensor:
        return self.U[: self.u43_bytes].view(torch.float16)

    def v43_halves(self) -> torch.Tensor:
        return self.V[: self.v43_bytes].view(torch.float16)

    def m43_halves(self) -> torch.Tensor:
        return self.M[: self.m43_bytes].view(torch.float16)


class WinogradConvFP8:
    """The FP8 compensated form of one problem (own output and FP8 workspaces).

    transform_filter holds the FP16 transformed filter (u16_bytes) followed by the FP8 filter
    records; transform_pin holds the FP8 input records."""

    def __init__(self, x: torch.Tensor, w: torch.Tensor, pad=(1, 1), out: torch.Tensor | None = None):
        assert x.dtype == torch.float16 and w.dtype == torch.float16
        assert x.is_cuda and w.is_cuda and x.is_contiguous() and w.is_contiguous()
        self.L = load_library()
        n, c, h, wd = x.shape
        k, c2, r, s = w.shape
        assert c == c2
        self.x, self.w = x, w
        self.shape = (n, c, h, wd, k, r, s, 1, 1, pad[0], pad[1])
        self.oh = (h - r + 2 * pad[0]) + 1
        self.ow = (wd - s + 2 * pad[1]) + 1
        probe = make_problem(*self.shape)
        self.v_bytes, self.u_bytes = fp8_workspace_size(probe)
        self.u16_bytes = form_workspace_size(probe, FORM_STAGED)[1]  # the FP16 (F(2x2)) U before the records
        dev = x.device
        self.out = out if out is not None else torch.empty((n, k, self.oh, self.ow), dtype=torch.float16, device=dev)
        self.V = torch.empty(max(self.v_bytes, 2), dtype=torch.uint8, device=dev)
        self.U = torch.empty(max(self.u_bytes, 2), dtype=torch.uint8, device=dev)
        self.pb = make_problem(*self.shape, ptrs={
            "in_": x.data_ptr(), "weight": w.data_ptr(), "out": self.out.data_ptr(),
            "transform_pin": self.V.data_ptr(), "transform_filter": self.U.data_ptr(), "mid_result": None})

    def prepare_weights(self, stream=None) -> None:
        check(self.L.conv2d_fp8_prepare_weights(ctypes.byref(self.pb), _stream_handle(stream)),
              "conv2d_fp8_prepare_weights")

    def input_transform(self, stream=None) -> None:
        check(self.L.conv2d_fp8_input_transform(ctypes.byref(self.pb), _stream_handle(stream)),
              "conv2d_fp8_input_transform")

    def gemm(self, cfg: int = -1, stream=None) -> None:
        check(self.L.conv2d_fp8_gemm(ctypes.byref(self.pb), cfg, _stream_handle(stream)), "conv2d_fp8_gemm")

    def forward(self, cfg: int = -1, stream=None) -> None:
        check(self.L.conv2d_fp8_forward(ctypes.byref(self.pb), cfg, _stream_handle(stream)), "conv2d_fp8_forward")

    def selected_config(self) -> int:
        c = ctypes.c_int()
        check(self.L.conv2d_fp8_select_tile_config(ctypes.byref(self.pb), ctypes.byref(c)),
              "conv2d_fp8_select_tile_config")
        return c.value

    def direct_f64(self, stream=None) -> torch.Tensor:
        ref = torch.empty(self.out.shape, dtype=torch.float64, device=self.x.device)
        check(self.L.conv2d_direct_f64(ctypes.byref(self.pb), ctypes.c_void_p(ref.data_ptr()), _stream_handle(stream)),
              "conv2d_direct_f64")
        return ref

    def u8_bytes(self) -> torch.Tensor:
        return self.U[self.u16_bytes: self.u_bytes]

    def u16_halves(self) -> torch.Tensor:
        return self.U[: self.u16_bytes].view(torch.float16)

    def v8_bytes(self) -> torch.Tensor:
        return self.V[: self.v_bytes]


def direct_f64(x: torch.Tensor, w: torch.Tensor, stride=(1, 1), pad=(1, 1)) -> torch.Tensor:
    """FP64 validation conv of any geometry through the library's GPU kernel."""
    L = load_library()
    n, c, h, wd = x.shape
    k, _, r, s = w.shape
    oh = (h - r + 2 * pad[0]) // stride[0] + 1
    ow = (wd - s + 2 * pad[1]) // stride[1] + 1
    pb = make_problem(n, c, h, wd, k, r, s, stride[0], stride[1], pad[0], pad[1],
                      ptrs={"in_": x.data_ptr(), "weight": w.data_ptr()})
    ref = torch.empty((n, k, oh, ow), dtype=torch.float64, device=x.device)
    check(L.conv2d_direct_f64(ctypes.byref(pb), ctypes.c_void_p(ref.data_ptr()), _stream_handle(None)),
          "conv2d_direct_f64")
    return ref


class Batch:
    """The multi-layer form over several WinogradConv problems (each keeps its own buffers)."""

    def __init__(self, convs: list[WinogradConv]):
        self.L = load_library()
        self.convs = convs
        arr = (type(convs[0].pb) * len(convs))(*[c.pb for c in convs])
        self._arr = arr
        h = ctypes.c_void_p()
        check(self.L.conv2d_batch_create(arr, len(convs), ctypes.byref(h)), "conv2d_batch_create")
        self.h = h

    def close(self) -> None:
        if self.h:
            check(self.L.conv2d_batch_destroy(self.h), "conv2d_batch_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def prepare_weights(self, stream=None):
        check(self.L.conv2d_batch_prepare_weights(self.h, _stream_handle(stream)), "conv2d_batch_prepare_weights")

    def forward(self, stream=None):
        check(self.L.conv2d_batch_forward(self.h, _stream_handle(stream)), "conv2d_batch_forward")

    def forward_separate(self, stream=None):
        check(self.L.conv2d_batch_forward_separate(self.h, _stream_handle(stream)), "conv2d_batch_forward_separate")

    def capture_graph(self, stream):
        check(self.L.conv2d_batch_capture_graph(self.h, _stream_handle(stream)), "conv2d_batch_capture_graph")

    def launch_graph(self, stream=None):
        check(self.L.conv2d_batch_launch_graph(self.h, _stream_handle(stream)), "conv2d_batch_launch_graph")

    def autotune(self, reps: int = 3, stream=None):
        check(self.L.conv2d_batch_autotune(self.h, reps, _stream_handle(stream)), "conv2d_batch_autotune")

    def layer_configs(self) -> list[int]:
        out = []
        for i in range(len(self.convs)):
            c = ctypes.c_int()
            check(self.L.conv2d_batch_layer_config(self.h, i, ctypes.byref(c)), "conv2d_batch_layer_config")
            out.append(c.value)
        return out

    def layer_forms(self) -> list[int]:
        out = []
        for i in range(len(self.convs)):
            f = ctypes.c_int()
            check(self.L.conv2d_batch_layer_form(self.h, i, ctypes.byref(f)), "conv2d_batch_layer_form")
            out.append(f.value)
        return out


def batch_layer_shapes(seed: int = 4, count: int = 24):
    """The 24-layer batched workload: seeded shapes N in {1,4,16}, C,K in
    {3,32,64,96,128,256,512}, H=W in {7,13,14,27,28,56,112}, 3x3 stride 1 pad 1."""
    import numpy as np
    rng = np.random.default_rng(seed)
    ns, cs, hs = [1, 4, 16], [3, 32, 64, 96, 128, 256, 512], [7, 13, 14, 27, 28, 56, 112]
    return [(int(rng.choice(ns)), int(rng.choice(cs)), int(rng.choice(hs)), int(rng.choice(cs)))
            for _ in range(count)]
