"""Seeded synthetic image pairs for the ssimulacra2 workloads.

Thin ctypes wrapper over libssim2synth.so (csrc/synth.c, host-only C). The
generators are byte-deterministic; tests pin SHA-256 digests of their output.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "libssim2synth.so"
if not _LIB_PATH.exists():
    raise ImportError(f"{_LIB_PATH} is missing: run __graft_entry__.build()")
_L = ctypes.CDLL(str(_LIB_PATH))
_L.synth_pair.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                          ctypes.c_void_p, ctypes.c_void_p]
_L.synth_pair.restype = ctypes.c_int
_L.synth_mix_seed.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
_L.synth_mix_seed.restype = ctypes.c_uint64
_L.synth_log_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
_L.synth_log_uniform.restype = ctypes.c_int
_L.synth_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
_L.synth_uniform.restype = ctypes.c_double
_u16p = ctypes.c_void_p
_L.synth_yuv420_pair.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [_u16p] * 6
_L.synth_yuv420_pair.restype = ctypes.c_int

NOISE_GRAD_BLOCKMEAN = 1   # 4 octaves value noise + gradient; Gaussian noise + 8x8 block means
NOISE_RECT_DCT = 2         # value noise + 5 rectangles; 8x8 DCT quantisation
NOISE_GRAD_DCT = 4         # noise + gradient reference, DCT quantisation


def pair(kind: int, seed: int, width: int, height: int, step: float = 16.0):
    """Return (ref, dis) as (h, w, 3) u8 arrays."""
    ref = np.empty((height, width, 3), np.uint8)
    dis = np.empty((height, width, 3), np.uint8)
    if _L.synth_pair(kind, seed & (2**64 - 1), width, height, float(step), ref.ctypes.data, dis.ctypes.data):
        raise ValueError("synth_pair: bad arguments")
    return ref, dis


def mix_seed(seed: int, index: int) -> int:
    return _L.synth_mix_seed(seed & (2**64 - 1), index)


def log_uniform(seed: int, index: int, salt: int, lo: int, hi: int) -> int:
    return _L.synth_log_uniform(seed & (2**64 - 1), index, salt, lo, hi)


def uniform(seed: int, index: int, salt: int) -> float:
    return _L.synth_uniform(seed & (2**64 - 1), index, salt)


def yuv420_pair(seed: int, width: int, height: int, frame: int):
    """10-bit BT.709 limited-range 4:2:0 frames: ((Y, Cb, Cr) ref, (Y, Cb, Cr) dis), u16."""
    cw, ch = (width + 1) // 2, (height + 1) // 2
    planes = [np.empty((height, width), np.uint16), np.empty((ch, cw), np.uint16), np.empty((ch, cw), np.uint16)]
    dplanes = [np.empty_like(p) for p in planes]
    if _L.synth_yuv420_pair(seed & (2**64 - 1), width, height, frame,
                            *[p.ctypes.data for p in planes + dplanes]):
        raise ValueError("synth_yuv420_pair: bad arguments")
    return tuple(planes), tuple(dplanes)
