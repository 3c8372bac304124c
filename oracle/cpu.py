"""ctypes binding of ``cpu_chain.c``: the all-core FP32 CPU restatement of
the benchmark workloads.  TEST INFRASTRUCTURE ONLY: the benchmark's CPU
baseline leg times it, and tests/ pin it against the FP64 oracle.

The colour matrices are folded here in FP64 from ``oracle.color`` (as a CPU
implementation folds them once) and handed to C as float32.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from . import color as C

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libzs_cpu.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            subprocess.run(["make", "-s", "-C", _HERE], check=True)
        l = ctypes.CDLL(_LIB_PATH)
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
        l.zsc_threads.restype = i32
        l.zsc_f2h.argtypes = [ctypes.c_float]
        l.zsc_f2h.restype = ctypes.c_uint16
        l.zsc_c3_image.argtypes = [vp, i32, i32, i32, vp, i32, i32, vp, vp]
        l.zsc_c2.argtypes = [vp, vp, vp, i64, vp, vp, vp]
        l.zsc_c4.argtypes = [vp, vp, i64, vp, vp, vp]
        l.zsc_c5.argtypes = [vp, i32, i32, i32, vp, vp, i32, i32, vp, vp, vp]
        l.zsc_c1.argtypes = [vp, vp, i32, i32, vp, vp, i32, i32]
        l.zsc_init()
        _lib = l
    return _lib


def threads() -> int:
    return lib().zsc_threads()


def _f32(m):
    return np.ascontiguousarray(np.asarray(m, dtype=np.float64).reshape(-1), dtype=np.float32)


def _p(a):
    return a.ctypes.data


def oklab_chain_matrices(primaries: int, pre=None, chroma: float = 1.0, hue_deg: float = 0.0):
    """(A, mid, C) of lms = A rgb, lms' = cbrt, mid, cube, rgb = C lms for an
    Oklab round trip in ``primaries`` with the (a, b) adjustment folded in;
    ``pre`` is a matrix applied to rgb first."""
    to_xyz = C.rgb_to_xyz_matrix(primaries, 5)
    A = C.OKLAB_M1 @ to_xyz
    if pre is not None:
        A = A @ np.asarray(pre, dtype=np.float64)
    t = np.radians(hue_deg)
    adj = np.array([[1.0, 0.0, 0.0], [0.0, chroma * np.cos(t), -chroma * np.sin(t)],
                    [0.0, chroma * np.sin(t), chroma * np.cos(t)]])
    mid = np.linalg.inv(C.OKLAB_M2) @ adj @ C.OKLAB_M2
    Cm = np.linalg.inv(to_xyz) @ np.linalg.inv(C.OKLAB_M1)
    return _f32(A), _f32(mid), _f32(Cm)


def _inv2x3(m):
    inv = np.linalg.inv(np.asarray(m, dtype=np.float32).astype(np.float64).reshape(3, 3))
    return np.ascontiguousarray(inv[:2].reshape(-1), dtype=np.float64)


def c2(src_u8, layer_f16_bits):
    """C2 over n pixels: (n, 4) uint8 and (n, 4) uint16 half bits -> (n, 4) uint8."""
    src = np.ascontiguousarray(src_u8, dtype=np.uint8).reshape(-1, 4)
    lay = np.ascontiguousarray(layer_f16_bits, dtype=np.uint16).reshape(-1, 4)
    out = np.empty_like(src)
    A, mid, Cm = oklab_chain_matrices(7, pre=C.convert_matrix(3, 7, 5), chroma=1.2)
    lib().zsc_c2(_p(src), _p(lay), _p(out), src.shape[0], _p(A), _p(mid), _p(Cm))
    return out


def c4(src_u8, hue_deg=30.0, chroma=1.2):
    src = np.ascontiguousarray(src_u8, dtype=np.uint8).reshape(-1, 4)
    out = np.empty_like(src)
    A, mid, Cm = oklab_chain_matrices(3, chroma=chroma, hue_deg=hue_deg)
    lib().zsc_c4(_p(src), _p(out), src.shape[0], _p(A), _p(mid), _p(Cm))
    return out


def c3_image(data, fmt, w, h, ow, oh):
    """One C3 image: bytes (h, w * bpp) -> bytes (oh, ow * bpp)."""
    src = np.ascontiguousarray(data)
    out = np.empty((oh, ow * (4 if fmt == 0 else 8)), dtype=np.uint8)
    m3 = _f32(C.convert_matrix(3, 7, 5))
    inv = _inv2x3(np.diag([np.float32(ow / w), np.float32(oh / h), 1.0]))
    lib().zsc_c3_image(_p(src), fmt, w, h, _p(out), ow, oh, _p(m3), _p(inv))
    return out


def c5(layers_f16_bits, matrices, w, h):
    """layers: list of (h, w * 8) uint8 (RGBA16F bytes); returns (h, w * 4) uint8."""
    ls = [np.ascontiguousarray(l) for l in layers_f16_bits]
    ptrs = (ctypes.c_void_p * len(ls))(*[l.ctypes.data for l in ls])
    inv = np.ascontiguousarray(np.concatenate([_inv2x3(m) for m in matrices]))
    out = np.empty((h, w * 4), dtype=np.uint8)
    A, mid, Cm = oklab_chain_matrices(3, chroma=1.2, hue_deg=30.0)
    lib().zsc_c5(ctypes.cast(ptrs, ctypes.c_void_p), len(ls), w, h, _p(inv), _p(out), w, h, _p(A), _p(mid), _p(Cm))
    return out


def c1(a_u8, b_u8, matrix, w, h):
    a = np.ascontiguousarray(a_u8, dtype=np.uint8)
    b = np.ascontiguousarray(b_u8, dtype=np.uint8)
    out = np.empty((h, w * 4), dtype=np.uint8)
    lib().zsc_c1(_p(a), _p(b), w, h, _p(_inv2x3(matrix)), _p(out), w, h)
    return out
