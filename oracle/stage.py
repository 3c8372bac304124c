"""ctypes binding of ``stage_frag.c`` (FP32 restatement of
lib/std/src/stage.frag).  TEST INFRASTRUCTURE ONLY."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libzs_oracle.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        u32, i64, fp = ctypes.c_uint32, ctypes.c_int64, ctypes.c_void_p
        _lib.zso_stage_decode_n.argtypes = [fp, i64, u32, u32, u32, fp]
        _lib.zso_stage_encode_n.argtypes = [fp, i64, u32, u32, u32, fp]
        _lib.zso_transfer_decode1.argtypes = [ctypes.c_float, u32]
        _lib.zso_transfer_decode1.restype = ctypes.c_float
        _lib.zso_transfer_encode1.argtypes = [ctypes.c_float, u32]
        _lib.zso_transfer_encode1.restype = ctypes.c_float
    return _lib


def stage_decode(nums, bits: int, parts: int, transfer: int) -> np.ndarray:
    """stage.frag decode (demux -> parts_normalize -> parts_untransfer) of
    u32 texel values; returns (N, 4) float32."""
    nums = np.ascontiguousarray(nums, dtype=np.uint32).ravel()
    out = np.empty((nums.size, 4), dtype=np.float32)
    lib().zso_stage_decode_n(nums.ctypes.data, nums.size, bits, parts, transfer, out.ctypes.data)
    return out


def stage_encode(values, bits: int, parts: int, transfer: int) -> np.ndarray:
    """stage.frag encode (parts_transfer -> parts_denormalize -> clamp ->
    truncating mux) of (N, 4) float32 linear values; returns u32 texels."""
    values = np.ascontiguousarray(values, dtype=np.float32).reshape(-1, 4)
    out = np.empty(values.shape[0], dtype=np.uint32)
    lib().zso_stage_encode_n(values.ctypes.data, values.shape[0], bits, parts, transfer, out.ctypes.data)
    return out
