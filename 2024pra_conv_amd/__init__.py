"""MI355X-native FP16 Winograd F(2x2,3x3) conv2d — Python bindings to the C-ABI library.

The product is the gfx950 HIP library ``lib/libwinoconv.so`` behind ``include/conv2d.h``
(the reference plugin interface, include/conv2d.h:8-70 of Yhorm26/2024PRA_CONV). This module
only mirrors that interface with ctypes so tests and benchmarks can drive it; device memory
and streams come from PyTorch (plumbing only). There is no CPU fallback: loading fails
loudly when the library is missing.

Because the directory name starts with a digit, import it with
``importlib.import_module("2024pra_conv_amd")``.
"""
from __future__ import annotations

import ctypes
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libwinoconv.so")
DEMO_PATH = os.path.join(PKG_DIR, "lib", "conv2dfp16demo")

# error codes (include/conv2d.h)
OK, ERR_NULL, ERR_UNSUPPORTED, ERR_OVERFLOW, ERR_WORKSPACE, ERR_HIP, ERR_INVALID, ERR_IO, ERR_MISMATCH = range(9)
METHOD_WINOGRAD, METHOD_DIRECT = 0, 1
ACT_NONE, ACT_RELU = 0, 1
FORM_STAGED, FORM_FUSED, FORM_F43 = 0, 1, 2
IO_NCHW, IO_IN_NHWC, IO_OUT_NHWC, IO_OUT_BF16 = 0, 1, 2, 4
ERROR_NAMES = {
    ERR_NULL: "CONV2D_ERR_NULL",
    ERR_UNSUPPORTED: "CONV2D_ERR_UNSUPPORTED",
    ERR_OVERFLOW: "CONV2D_ERR_OVERFLOW",
    ERR_WORKSPACE: "CONV2D_ERR_WORKSPACE",
    ERR_HIP: "CONV2D_ERR_HIP",
    ERR_INVALID: "CONV2D_ERR_INVALID",
    ERR_IO: "CONV2D_ERR_IO",
    ERR_MISMATCH: "CONV2D_ERR_MISMATCH",
}


class Conv2dError(RuntimeError):
    def __init__(self, code: int, where: str):
        super().__init__(f"{where} returned {ERROR_NAMES.get(code, code)}")
        self.code = code


class Problem(ctypes.Structure):
    """problem_t (reference include/conv2d.h:8-27): six device pointers, eleven unsigned."""
    _fields_ = [
        ("in_", ctypes.c_void_p),
        ("weight", ctypes.c_void_p),
        ("out", ctypes.c_void_p),
        ("transform_pin", ctypes.c_void_p),
        ("transform_filter", ctypes.c_void_p),
        ("mid_result", ctypes.c_void_p),
    ] + [(f, ctypes.c_uint) for f in ("n", "c", "h", "w", "k", "r", "s", "u", "v", "p", "q")]


def _slot_fields(i: int):
    return [(f"blockx{i}", ctypes.c_uint), (f"blocky{i}", ctypes.c_uint), (f"blockz{i}", ctypes.c_uint),
            (f"threadx{i}", ctypes.c_uint), (f"thready{i}", ctypes.c_uint), (f"threadz{i}", ctypes.c_uint),
            (f"dynmicLdsSize{i}", ctypes.c_uint), (f"kernelPtr{i}", ctypes.c_void_p)]


class KernelInfo(ctypes.Structure):
    """kernelInfo_t (reference include/conv2d.h:29-66): four launch slots."""
    _fields_ = _slot_fields(1) + _slot_fields(2) + _slot_fields(3) + _slot_fields(4)

    def slot(self, i: int) -> dict:
        g = lambda f: getattr(self, f"{f}{i}")
        return {"grid": (g("blockx"), g("blocky"), g("blockz")),
                "block": (g("threadx"), g("thready"), g("threadz")),
                "lds": g("dynmicLdsSize"), "kernel": g("kernelPtr")}


class ParamType(ctypes.Structure):
    """mykernelParamType (include/mykernelParamType.h): reference fields + appended geometry."""
    _fields_ = [(f, ctypes.c_void_p) for f in
                ("pin", "pweight", "pout", "transform_pin", "transform_filter", "mid_result")] + \
               [(f, ctypes.c_short) for f in
                ("n", "c", "h", "w", "k", "r", "s", "u", "v", "p", "q", "Oh", "Ow")] + \
               [(f, ctypes.c_int) for f in
                ("tiles_h", "tiles_w", "T", "NT", "G32", "CB", "KB32", "band", "cfg", "flags", "nsteps",
                 "fbh", "fbw", "fnbh", "fnbw", "fpr", "fpc", "fsh", "fnu", "act", "form")] + \
               [("bias", ctypes.c_void_p), ("io", ctypes.c_int), ("f43_npos", ctypes.c_int)]


_lib = None

# exported symbols and their ctypes signatures (every function declared in include/conv2d.h)
_P = ctypes.POINTER
_SIGS = {
    "getParamsize": ([_P(Problem), _P(ctypes.c_int)], ctypes.c_int),
    "getkernelInfo": ([_P(Problem), _P(KernelInfo), ctypes.c_void_p], ctypes.c_int),
    "getkernelInfo_io": ([_P(Problem), ctypes.c_int, _P(KernelInfo), ctypes.c_void_p], ctypes.c_int),
    "conv2d_forward_io": ([_P(Problem), ctypes.c_int, ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    "conv2d_direct_f64_io": ([_P(Problem), ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "conv2d_workspace_size": ([_P(Problem), _P(ctypes.c_size_t), _P(ctypes.c_size_t), _P(ctypes.c_size_t)], ctypes.c_int),
    "conv2d_launch": ([_P(KernelInfo), ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "conv2d_prepare_weights": ([_P(Problem), ctypes.c_void_p], ctypes.c_int),
    "conv2d_input_transform": ([_P(Problem), ctypes.c_void_p], ctypes.c_int),
    "conv2d_gemm": ([_P(Problem), ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    "conv2d_forward": ([_P(Problem), ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    "conv2d_fused_gemm": ([_P(Problem), ctypes.c_void_p], ctypes.c_int),
    "conv2d_form_workspace_size": ([_P(Problem), ctypes.c_int, _P(ctypes.c_size_t), _P(ctypes.c_size_t),
                                    _P(ctypes.c_size_t)], ctypes.c_int),
    "conv2d_f43_prepare_weights": ([_P(Problem), ctypes.c_void_p], ctypes.c_int),
    "conv2d_f43_input_transform": ([_P(Problem), ctypes.c_void_p], ctypes.c_int),
    "conv2d_f43_gemm": ([_P(Problem), ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    "conv2d_f43_output_transform": ([_P(Problem), ctypes.c_void_p], ctypes.c_int),
    "conv2d_f43_forward": ([_P(Problem), ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    "conv2d_f43_num_tile_configs": ([], ctypes.c_int),
    "conv2d_f43_tile_config": ([ctypes.c_int, _P(ctypes.c_int), _P(ctypes.c_int), _P(ctypes.c_int)], ctypes.c_int),
    "conv2d_f43_select_tile_config": ([_P(Problem), _P(ctypes.c_int)], ctypes.c_int),
    "conv2d_forward_static": ([_P(Problem), ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    "conv2d_static_scratch_bytes": ([], ctypes.c_size_t),
    "conv2d_form": ([_P(Problem), _P(ctypes.c_int)], ctypes.c_int),
    "conv2d_weights_save": ([_P(Problem), ctypes.c_char_p, ctypes.c_void_p], ctypes.c_int),
    "conv2d_weights_load": ([_P(Problem), ctypes.c_char_p, ctypes.c_void_p], ctypes.c_int),
    "conv2d_weight_hash": ([ctypes.c_void_p, ctypes.c_size_t, _P(ctypes.c_ulonglong)], ctypes.c_int),
    "conv2d_set_epilogue": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int], ctypes.c_int),
    "conv2d_forward_epilogue": ([_P(Problem), ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    "conv2d_num_tile_configs": ([], ctypes.c_int),
    "conv2d_tile_config": ([ctypes.c_int, _P(ctypes.c_int), _P(ctypes.c_int), _P(ctypes.c_int)], ctypes.c_int),
    "conv2d_select_tile_config": ([_P(Problem), _P(ctypes.c_int)], ctypes.c_int),
    "conv2d_direct_f64": ([_P(Problem), ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "conv2d_method": ([_P(Problem), _P(ctypes.c_int)], ctypes.c_int),
    "conv2d_direct_num_tile_configs": ([], ctypes.c_int),
    "conv2d_direct_tile_config": ([ctypes.c_int, _P(ctypes.c_int), _P(ctypes.c_int), _P(ctypes.c_int)], ctypes.c_int),
    "conv2d_fp8_workspace_size": ([_P(Problem), _P(ctypes.c_size_t), _P(ctypes.c_size_t)], ctypes.c_int),
    "conv2d_fp8_prepare_weights": ([_P(Problem), ctypes.c_void_p], ctypes.c_int),
    "conv2d_fp8_input_transform": ([_P(Problem), ctypes.c_void_p], ctypes.c_int),
    "conv2d_fp8_gemm": ([_P(Problem), ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    "conv2d_fp8_forward": ([_P(Problem), ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    "conv2d_fp8_num_tile_configs": ([], ctypes.c_int),
    "conv2d_fp8_tile_config": ([ctypes.c_int, _P(ctypes.c_int), _P(ctypes.c_int), _P(ctypes.c_int)], ctypes.c_int),
    "conv2d_fp8_select_tile_config": ([_P(Problem), _P(ctypes.c_int)], ctypes.c_int),
    "conv2d_tuned_set": ([_P(Problem), ctypes.c_int], ctypes.c_int),
    "conv2d_tuned_set_form": ([_P(Problem), ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "conv2d_tuned_clear": ([], ctypes.c_int),
    "conv2d_tuned_count": ([], ctypes.c_int),
    "conv2d_tuned_save": ([ctypes.c_char_p], ctypes.c_int),
    "conv2d_tuned_load": ([ctypes.c_char_p], ctypes.c_int),
    "conv2d_autotune": ([_P(Problem), ctypes.c_int, ctypes.c_void_p, _P(ctypes.c_int), _P(ctypes.c_float),
                         ctypes.c_int], ctypes.c_int),
    "conv2d_batch_create": ([_P(Problem), ctypes.c_int, _P(ctypes.c_void_p)], ctypes.c_int),
    "conv2d_batch_destroy": ([ctypes.c_void_p], ctypes.c_int),
    "conv2d_batch_prepare_weights": ([ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "conv2d_batch_forward": ([ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "conv2d_batch_forward_separate": ([ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "conv2d_batch_capture_graph": ([ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "conv2d_batch_launch_graph": ([ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "conv2d_batch_autotune": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    "conv2d_batch_layer_config": ([ctypes.c_void_p, ctypes.c_int, _P(ctypes.c_int)], ctypes.c_int),
"conv2d_batch_layer_form": ([ctypes.c_void_p, ctypes.c_int, _P(ctypes.c_int)], ctypes.c_int),
    "conv2d_mlp_param_count": ([], ctypes.c_int),
    "conv2d_mlp_features": ([_P(Problem), ctypes.c_int, _P(ctypes.c_float)], ctypes.c_int),
    "conv2d_mlp_forward": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p],
                           ctypes.c_int),
    "conv2d_mlp_loss_grad": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                              ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "conv2d_mlp_adam_step": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                              ctypes.c_float, ctypes.c_void_p], ctypes.c_int),
    "conv2d_mlp_train": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                          ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float, _P(ctypes.c_float),
                          ctypes.c_void_p], ctypes.c_int),
    "conv2d_mlp_select": ([ctypes.c_void_p, _P(Problem), _P(ctypes.c_int), _P(ctypes.c_float), ctypes.c_void_p],
                          ctypes.c_int),
    "conv2d_mlp_select_form": ([ctypes.c_void_p, ctypes.c_void_p, _P(Problem), _P(ctypes.c_int), _P(ctypes.c_int),
                                _P(ctypes.c_float), ctypes.c_void_p], ctypes.c_int),
    "conv2d_version": ([], ctypes.c_char_p),
}


def load_library(path: str | None = None):
    """Loads libwinoconv.so (raises if it is not built: there is no fallback). WINOCONV_LIB
    names another build of the same library (the sanitizer build of build_lib.py)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or os.environ.get("WINOCONV_LIB") or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is not built; run __graft_entry__.build() (no CPU fallback exists)")
    L = ctypes.CDLL(p)
    for name, (args, res) in _SIGS.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    if path is None:
        _lib = L
    return L


def check(rc: int, where: str) -> None:
    if rc != OK:
        raise Conv2dError(rc, where)


def make_problem(n, c, h, w, k, r=3, s=3, u=1, v=1, p=1, q=1, ptrs=None) -> Problem:
    pb = Problem()
    for f, val in zip(("n", "c", "h", "w", "k", "r", "s", "u", "v", "p", "q"), (n, c, h, w, k, r, s, u, v, p, q)):
        setattr(pb, f, val)
    if ptrs:
        for f, val in ptrs.items():
            setattr(pb, f, val)
    return pb


def workspace_size(pb: Problem):
    L = load_library()
    a, b, c = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
    check(L.conv2d_workspace_size(ctypes.byref(pb), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
          "conv2d_workspace_size")
    return a.value, b.value, c.value


def form_workspace_size(pb: Problem, form: int):
    """(V, U, M) bytes of one form, or None where that form does not apply."""
    L = load_library()
    a, b, c = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
    rc = L.conv2d_form_workspace_size(ctypes.byref(pb), form, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    if rc == ERR_UNSUPPORTED:
        return None
    check(rc, "conv2d_form_workspace_size")
    return a.value, b.value, c.value


def tile_configs():
    L = load_library()
    out = []
    for i in range(L.conv2d_num_tile_configs()):
        bk, bt, th = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(L.conv2d_tile_config(i, ctypes.byref(bk), ctypes.byref(bt), ctypes.byref(th)), "conv2d_tile_config")
        out.append({"cfg": i, "bk": bk.value, "bt": bt.value, "threads": th.value})
    return out


def fp8_workspace_size(pb: Problem):
    L = load_library()
    a, b = ctypes.c_size_t(), ctypes.c_size_t()
    check(L.conv2d_fp8_workspace_size(ctypes.byref(pb), ctypes.byref(a), ctypes.byref(b)), "conv2d_fp8_workspace_size")
    return a.value, b.value


def fp8_tile_configs():
    L = load_library()
    out = []
    for i in range(L.conv2d_fp8_num_tile_configs()):
        bk, bt, th = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(L.conv2d_fp8_tile_config(i, ctypes.byref(bk), ctypes.byref(bt), ctypes.byref(th)),
              "conv2d_fp8_tile_config")
        out.append({"cfg": i, "bk": bk.value, "bt": bt.value, "threads": th.value})
    return out


def exported_symbols():
    return list(_SIGS)
