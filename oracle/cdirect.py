"""TEST INFRASTRUCTURE ONLY — ctypes bindings for oracle/direct_conv.c and oracle/_ref.

  conv2d_f64(x16, w16, geom)          exact FP64 direct conv (verfiy.h:5-43, exact products)
  conv2d_f64_threads(...)             same, bit-identical, split over threads by k
  conv2dcpu_faithful(x16, w16, geom)  the reference checker's arithmetic restated
                                      (FP16-rounded products, FP64 sum, FP16 out; verfiy.h:32,38)
  ref_conv2dcpu(x16, w16, geom)       the reference's own conv2dcpu compiled from its source
                                      (None when oracle/_ref is not built)
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(HERE, "_build", "liboracle.so")
_REF = os.path.join(HERE, "_ref", "libref_conv.so")
_lock = threading.Lock()
_lib = None
_ref = None


def build() -> None:
    """Compiles the C restatement (and the reference checker when its source is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if os.path.exists("/root/reference/include/verfiy.h"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(_LIB):
                subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
            L = ctypes.CDLL(_LIB)
            i = ctypes.c_int
            p = ctypes.c_void_p
            L.oracle_conv2d_f64.argtypes = [p, p, p] + [i] * 11
            L.oracle_conv2d_f64_decoded.argtypes = [p, p, p] + [i] * 13
            L.oracle_conv2dcpu_faithful.argtypes = [p, p, p] + [i] * 11
            L.oracle_decode_halves.argtypes = [p, p, ctypes.c_size_t]
            L.oracle_round_half.argtypes = [ctypes.c_double]
            L.oracle_round_half.restype = ctypes.c_double
            _lib = L
        return _lib


def ref_lib():
    """The reference's own conv2dcpu, or None when oracle/_ref was not built."""
    global _ref
    with _lock:
        if _ref is None and os.path.exists(_REF):
            L = ctypes.CDLL(_REF)
            i = ctypes.c_int
            p = ctypes.c_void_p
            L.ref_conv2dcpu.argtypes = [p, p, p] + [i] * 11
            _ref = L
        return _ref


def _dims(x16: np.ndarray, w16: np.ndarray, stride=(1, 1), pad=(1, 1)):
    n, c, h, w = x16.shape
    k, c2, r, s = w16.shape
    assert c == c2
    u, v = stride
    p, q = pad
    oh = (h - r + 2 * p) // u + 1
    ow = (w - s + 2 * q) // v + 1
    return (n, c, h, w, k, r, s, u, v, p, q), (n, k, oh, ow)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def conv2d_f64(x16, w16, stride=(1, 1), pad=(1, 1)) -> np.ndarray:
    x16 = np.ascontiguousarray(x16, dtype=np.float16)
    w16 = np.ascontiguousarray(w16, dtype=np.float16)
    d, oshape = _dims(x16, w16, stride, pad)
    out = np.empty(oshape, dtype=np.float64)
    lib().oracle_conv2d_f64(_ptr(x16), _ptr(w16), _ptr(out), *d)
    return out


def conv2d_f64_threads(x16, w16, stride=(1, 1), pad=(1, 1), threads: int = 8) -> np.ndarray:
    """Bit-identical to conv2d_f64; splits the k range over threads (ctypes drops the GIL)."""
    x16 = np.ascontiguousarray(x16, dtype=np.float16)
    w16 = np.ascontiguousarray(w16, dtype=np.float16)
    d, oshape = _dims(x16, w16, stride, pad)
    L = lib()
    din = np.empty(x16.shape, dtype=np.float64)
    dw = np.empty(w16.shape, dtype=np.float64)
    L.oracle_decode_halves(_ptr(x16), _ptr(din), x16.size)
    L.oracle_decode_halves(_ptr(w16), _ptr(dw), w16.size)
    out = np.empty(oshape, dtype=np.float64)
    k = d[4]
    bounds = np.linspace(0, k, min(threads, k) + 1).astype(int)
    ts = [threading.Thread(target=L.oracle_conv2d_f64_decoded,
                           args=(_ptr(din), _ptr(dw), _ptr(out), *d, int(a), int(b)))
          for a, b in zip(bounds[:-1], bounds[1:])]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    return out


def conv2dcpu_faithful(x16, w16, stride=(1, 1), pad=(1, 1)) -> np.ndarray:
    x16 = np.ascontiguousarray(x16, dtype=np.float16)
    w16 = np.ascontiguousarray(w16, dtype=np.float16)
    d, oshape = _dims(x16, w16, stride, pad)
    out = np.empty(oshape, dtype=np.float16)
    lib().oracle_conv2dcpu_faithful(_ptr(x16), _ptr(w16), _ptr(out), *d)
    return out


def ref_conv2dcpu(x16, w16, stride=(1, 1), pad=(1, 1)):
    L = ref_lib()
    if L is None:
        return None
    x16 = np.ascontiguousarray(x16, dtype=np.float16)
    w16 = np.ascontiguousarray(w16, dtype=np.float16)
    d, oshape = _dims(x16, w16, stride, pad)
    out = np.zeros(oshape, dtype=np.float16)
    L.ref_conv2dcpu(_ptr(x16), _ptr(w16), _ptr(out), *d)
    return out


def round_half(x: float) -> float:
    return lib().oracle_round_half(float(x))
