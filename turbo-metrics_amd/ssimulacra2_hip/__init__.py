"""ssimulacra2 on MI355X: Python host mirror of the reference's Rust API.

`Ssimulacra2` mirrors `ssimulacra2_cuda::Ssimulacra2`
(crates/ssimulacra2-cuda/src/lib.rs:27-602) method for method on top of the C
ABI in include/ssim2hip.h (libssim2hip.so, hand-written HIP kernels for
gfx950). Images are device buffers (`DeviceImage`, the counterpart of
cudarse-npp `Image<S, C<3>>`); streams are `Stream` objects or None for the
default stream.

The shared library must be built in-tree (``__graft_entry__.build()``); there is
no CPU fallback, and importing this package fails loudly when it is missing.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
# SSIM2HIP_LIBRARY may name another build of the same library (an installed
# copy, a differently tuned build).
LIB_PATH = Path(os.environ.get("SSIM2HIP_LIBRARY") or (_HERE / "libssim2hip.so"))

__all__ = [
    "Ssimulacra2",
    "DeviceImage",
    "Stream",
    "Ssim2Error",
    "lib",
    "LIB_PATH",
    "STATUS",
    "Ssim2Batch",
    "Ssim2Server",
    "STORAGE",
]

STORAGE = {"f32": 0, "f16": 1, "bf16": 2}
TILE_SHAPES = [(32, 16), (64, 16), (32, 32), (32, 24), (64, 32)]

STATUS = {
    0: "SSIM2_OK",
    1: "SSIM2_INVALID_SIZE",
    2: "SSIM2_DIM_MISMATCH",
    3: "SSIM2_NO_REFERENCE",
    4: "SSIM2_OOM",
    5: "SSIM2_BATCH_SIZE",
    6: "SSIM2_ADMISSION_REFUSED",
    7: "SSIM2_HIP_ERROR",
    8: "SSIM2_INVALID_ARGUMENT",
    9: "SSIM2_NOT_READY",
}


class Ssim2Error(RuntimeError):
    """A non-OK ssim2_status; `.status` holds the name, e.g. "SSIM2_DIM_MISMATCH"."""

    def __init__(self, code: int, message: str):
        self.code = code
        self.status = STATUS.get(code, f"SSIM2_{code}")
        super().__init__(f"{self.status}: {message}")


class _Image(ctypes.Structure):
    _fields_ = [
        ("ptr", ctypes.c_void_p),
        ("pitch_bytes", ctypes.c_size_t),
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
    ]


class _Pair(ctypes.Structure):
    _fields_ = [("ref", _Image), ("dis", _Image)]


class _RefPair(ctypes.Structure):
    _fields_ = [("ref_id", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("dis", _Image)]


class _Yuv(ctypes.Structure):
    _fields_ = [
        ("y", ctypes.c_void_p),
        ("c0", ctypes.c_void_p),
        ("c1", ctypes.c_void_p),
        ("y_pitch_bytes", ctypes.c_size_t),
        ("c_pitch_bytes", ctypes.c_size_t),
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("format", ctypes.c_uint32),
        ("reserved0", ctypes.c_uint32),
    ]


class _YuvPair(ctypes.Structure):
    _fields_ = [("ref", _Yuv), ("dis", _Yuv)]


YUV_FORMATS = {"yuv420p10": 0, "p016": 1}


class _Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("min", "max", "mean", "var", "sample_var", "stddev",
                                                 "sample_stddev", "p1", "p5", "p50", "p95", "p99")]


class _Frame(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("reserved0", ctypes.c_uint32), ("srgb", _Image), ("yuv", _Yuv)]


class _VideoOptions(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("struct_size", "every", "skip", "skip_ref", "skip_dis", "frames",
                                                 "mode", "storage", "batch", "reserved0")]


FRAME_SOURCE = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(_Frame))


class _BatchOptions(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32),
        ("storage", ctypes.c_uint32),
        ("max_pairs", ctypes.c_uint32),
        ("reserved0", ctypes.c_uint32),
        ("max_pixels", ctypes.c_uint64),
        ("memory_budget", ctypes.c_uint64),
    ]


class _ServerOptions(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32),
        ("storage", ctypes.c_uint32),
        ("slots", ctypes.c_uint32),
        ("max_pairs_per_batch", ctypes.c_uint32),
        ("max_pixels_per_batch", ctypes.c_uint64),
        ("hbm_budget", ctypes.c_uint64),
        ("tile_cfg", ctypes.c_uint32),
        ("grid_mult", ctypes.c_uint32),
    ]


class _Result(ctypes.Structure):
    _fields_ = [
        ("request_id", ctypes.c_uint64),
        ("score", ctypes.c_double),
        ("submit_time_s", ctypes.c_double),
        ("done_time_s", ctypes.c_double),
        ("status", ctypes.c_int),
        ("reserved0", ctypes.c_uint32),
    ]


class _ServerStats(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32),
        ("reserved0", ctypes.c_uint32),
        ("submitted", ctypes.c_uint64),
        ("completed", ctypes.c_uint64),
        ("refused", ctypes.c_uint64),
        ("batches", ctypes.c_uint64),
        ("pixels_completed", ctypes.c_uint64),
        ("waiting", ctypes.c_uint64),
        ("in_flight_bytes", ctypes.c_uint64),
        ("first_submit_s", ctypes.c_double),
        ("last_done_s", ctypes.c_double),
    ]


class _Options(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32),
        ("reserved0", ctypes.c_uint32),
        ("memory_budget", ctypes.c_uint64),
    ]


def _load():
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP extension first (python -c "
            "'import __graft_entry__ as g; g.build()')"
        )
    L = ctypes.CDLL(str(LIB_PATH))
    P = ctypes.POINTER
    vp, u32, u64, sz, dbl = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_double
    img = P(_Image)
    sig = {
        "ssim2_create": (u32, u32, P(_Options), P(vp)),
        "ssim2_destroy": (vp,),
        "ssim2_mem_usage": (vp, P(sz)),
        "ssim2_compute_linear_async": (vp, img, img, vp),
        "ssim2_get_score": (vp, P(dbl)),
        "ssim2_compute_linear_sync": (vp, img, img, vp, P(dbl)),
        "ssim2_compute_srgb_sync": (vp, img, img, img, img, vp, P(dbl)),
        "ssim2_compute_srgb_async": (vp, img, img, vp),
        "ssim2_compute_srgb_noalloc_async": (img, img, vp, vp),
        "ssim2_compute_from_host_srgb": (vp, vp, vp, sz, vp, P(dbl)),
        "ssim2_prepare_reference_from_host_srgb": (vp, vp, sz, vp),
        "ssim2_compute_from_host_dis_srgb": (vp, vp, sz, vp, P(dbl)),
        "ssim2_set_reference_linear": (vp, img, vp),
        "ssim2_set_reference_srgb": (vp, img, vp),
        "ssim2_compute_with_reference_linear": (vp, img, img, vp, P(dbl)),
        "ssim2_compute_with_reference_srgb": (vp, img, vp, P(dbl)),
        "ssim2_clear_reference": (vp,),
        "ssim2_has_reference": (vp, P(ctypes.c_int)),
        "ssim2_get_sums": (vp, P(dbl), P(ctypes.c_int), P(ctypes.c_int)),
        "ssim2_get_plane": (vp, u32, u32, u32, vp, sz),
        "ssim2_srgb_to_linear": (img, img, vp),
        "ssim2_downscale_by_2": (img, img, vp),
        "ssim2_linear_to_xyb": (img, img, vp),
        "ssim2_yuv420_to_linear": (P(_Yuv), img, vp),
        "ssim2_compute_yuv420_async": (vp, P(_Yuv), P(_Yuv), vp),
        "ssim2_compute_yuv420_sync": (vp, P(_Yuv), P(_Yuv), vp, P(dbl)),
        "ssim2_yuv_alloc": (u32, u32, u32, P(_Yuv)),
        "ssim2_yuv_free": (P(_Yuv),),
        "ssim2_yuv_upload": (P(_Yuv), vp, vp, vp),
        "ssim2_batch_compute_yuv_async": (vp, P(_YuvPair), u32, vp),
        "ssim2_stats_compute": (P(dbl), sz, P(_Stats)),
        "ssim2_video_compute_all": (u32, u32, P(_VideoOptions), FRAME_SOURCE, vp, FRAME_SOURCE, vp, vp, P(dbl), sz,
                                    P(sz), P(_Stats)),
        "ssim2_image_alloc": (u32, u32, u32, img),
        "ssim2_image_free": (img,),
        "ssim2_image_upload": (img, vp, sz),
        "ssim2_image_download": (img, vp, sz),
        "ssim2_stream_create": (P(vp),),
        "ssim2_stream_destroy": (vp,),
        "ssim2_stream_synchronize": (vp,),
        "ssim2_time_kernels": (vp, img, img, vp, ctypes.c_int, P(ctypes.c_float)),
        "ssim2_device_alloc": (sz, P(vp)),
        "ssim2_device_free": (vp,),
        "ssim2_device_copy": (vp, vp, sz, vp),
        "ssim2_device_synchronize": (),
        "ssim2_check_arithmetic": (u32, u32, u32, P(u64)),
        "ssim2_batch_create": (P(_BatchOptions), P(vp)),
        "ssim2_batch_destroy": (vp,),
        "ssim2_batch_mem_usage": (vp, P(sz)),
        "ssim2_batch_compute_async": (vp, P(_Pair), u32, vp),
        "ssim2_batch_add_reference": (vp, img, vp, P(u32)),
        "ssim2_batch_remove_reference": (vp, u32),
        "ssim2_batch_compute_ref_async": (vp, P(_RefPair), u32, vp),
        "ssim2_batch_get_scores": (vp, P(dbl), u32),
        "ssim2_batch_compute_sync": (vp, P(_Pair), u32, vp, P(dbl)),
        "ssim2_batch_query": (vp, P(ctypes.c_int)),
        "ssim2_batch_get_sums": (vp, u32, P(dbl), P(ctypes.c_int)),
        "ssim2_batch_set_tiling": (vp, u32, u32),
        "ssim2_batch_get_tiling": (vp, P(u32), P(u32)),
        "ssim2_batch_autotune": (vp, P(_Pair), u32, vp, u32, ctypes.c_char_p, sz),
        "ssim2_server_create": (P(_ServerOptions), P(vp)),
        "ssim2_server_destroy": (vp,),
        "ssim2_server_submit": (vp, P(_Pair), u64),
        "ssim2_server_add_reference": (vp, img, P(u32)),
        "ssim2_server_remove_reference": (vp, u32),
        "ssim2_server_submit_ref": (vp, u32, img, u64),
        "ssim2_server_poll": (vp, P(_Result), u32, P(u32)),
        "ssim2_server_drain": (vp, dbl),
        "ssim2_server_stats": (vp, P(_ServerStats)),
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = list(args)
        f.restype = ctypes.c_int
    L.ssim2_status_string.argtypes = [ctypes.c_int]
    L.ssim2_status_string.restype = ctypes.c_char_p
    L.ssim2_last_error.argtypes = []
    L.ssim2_last_error.restype = ctypes.c_char_p
    L.ssim2_version.argtypes = []
    L.ssim2_version.restype = ctypes.c_int
    return L


lib = _load()


def _check(code: int):
    if code != 0:
        raise Ssim2Error(code, lib.ssim2_last_error().decode(errors="replace"))


def _stream_ptr(stream):
    if stream is None:
        return None
    return stream.handle


class Stream:
    """A non-blocking hipStream_t (cudarse-driver CuStream, stream.rs:46)."""

    def __init__(self):
        h = ctypes.c_void_p()
        _check(lib.ssim2_stream_create(ctypes.byref(h)))
        self.handle = h

    def sync(self):
        _check(lib.ssim2_stream_synchronize(self.handle))

    def close(self):
        if self.handle:
            lib.ssim2_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceImage:
    """Packed-RGB device image, u8 or f32 (cudarse-npp Image<S, C<3>>)."""

    def __init__(self, width: int, height: int, dtype=np.uint8):
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.dtype(np.uint8), np.dtype(np.float32)):
            raise ValueError("DeviceImage holds u8 or f32 samples")
        self._img = _Image()
        _check(lib.ssim2_image_alloc(width, height, self.dtype.itemsize, ctypes.byref(self._img)))

    @property
    def width(self):
        return self._img.width

    @property
    def height(self):
        return self._img.height

    @property
    def pitch_bytes(self):
        return self._img.pitch_bytes

    @property
    def ptr(self):
        return self._img.ptr

    def as_struct(self):
        return ctypes.byref(self._img)

    @classmethod
    def from_host(cls, array: np.ndarray) -> "DeviceImage":
        """Upload an (h, w, 3) u8 or f32 array (copy_from_cpu, image/mod.rs:410-439)."""
        a = np.ascontiguousarray(array)
        if a.ndim != 3 or a.shape[2] != 3:
            raise ValueError("expected an (h, w, 3) array")
        im = cls(a.shape[1], a.shape[0], a.dtype)
        im.upload(a)
        return im

    def upload(self, array: np.ndarray):
        a = np.ascontiguousarray(array, dtype=self.dtype)
        if a.shape != (self.height, self.width, 3):
            raise ValueError(f"shape {a.shape} != {(self.height, self.width, 3)}")
        _check(lib.ssim2_image_upload(ctypes.byref(self._img), a.ctypes.data, a.strides[0]))

    def download(self) -> np.ndarray:
        out = np.empty((self.height, self.width, 3), dtype=self.dtype)
        _check(lib.ssim2_image_download(ctypes.byref(self._img), out.ctypes.data, out.strides[0]))
        return out

    def free(self):
        if self._img.ptr:
            lib.ssim2_image_free(ctypes.byref(self._img))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceYuv:
    """YUV 4:2:0 frame in device memory, BT.709 limited range, u16 samples
    (the decoded surfaces of crates/cudarse/cudarse-video/src/dec.rs).
    format "yuv420p10": planar Y, Cb, Cr with 10-bit LSB-aligned samples;
    "p016": Y plus interleaved CbCr with samples in the high bits."""

    def __init__(self, width: int, height: int, fmt: str = "yuv420p10"):
        self.fmt = fmt
        self._img = _Yuv()
        _check(lib.ssim2_yuv_alloc(width, height, YUV_FORMATS[fmt], ctypes.byref(self._img)))

    @property
    def width(self):
        return self._img.width

    @property
    def height(self):
        return self._img.height

    def as_struct(self):
        return ctypes.byref(self._img)

    @classmethod
    def from_planes(cls, y: np.ndarray, cb: np.ndarray, cr: np.ndarray, fmt: str = "yuv420p10") -> "DeviceYuv":
        """Upload 10-bit planes (u16, LSB-aligned). For "p016" the samples are
        shifted to the high bits and Cb/Cr interleaved on the host."""
        y = np.ascontiguousarray(y, dtype=np.uint16)
        h, w = y.shape
        cw, ch = (w + 1) // 2, (h + 1) // 2
        cb = np.ascontiguousarray(cb, dtype=np.uint16)
        cr = np.ascontiguousarray(cr, dtype=np.uint16)
        if cb.shape != (ch, cw) or cr.shape != (ch, cw):
            raise ValueError(f"chroma planes must be {(ch, cw)}")
        im = cls(w, h, fmt)
        if fmt == "p016":
            yy = np.ascontiguousarray(y << 6)
            uv = np.ascontiguousarray(np.stack([cb << 6, cr << 6], axis=-1))
            _check(lib.ssim2_yuv_upload(ctypes.byref(im._img), yy.ctypes.data, uv.ctypes.data, None))
        else:
            _check(lib.ssim2_yuv_upload(ctypes.byref(im._img), y.ctypes.data, cb.ctypes.data, cr.ctypes.data))
        return im

    def free(self):
        if self._img.y:
            lib.ssim2_yuv_free(ctypes.byref(self._img))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def yuv420_to_linear(src: DeviceYuv, dst: DeviceImage, stream: Stream | None = None):
    """CudaColorspace::biplanaryuv420_to_linearrgb_16 (crates/cuda-colorspace/src/lib.rs:79-123)."""
    _check(lib.ssim2_yuv420_to_linear(src.as_struct(), dst.as_struct(), _stream_ptr(stream)))


def device_synchronize():
    """hipDeviceSynchronize."""
    _check(lib.ssim2_device_synchronize())


class DeviceBuffer:
    """Raw device allocation (bytes), with the calibration copy kernel."""

    def __init__(self, nbytes: int):
        p = ctypes.c_void_p()
        _check(lib.ssim2_device_alloc(nbytes, ctypes.byref(p)))
        self.ptr, self.nbytes = p, nbytes

    def copy_from(self, other: "DeviceBuffer", nbytes: int | None = None, stream=None):
        _check(lib.ssim2_device_copy(self.ptr, other.ptr, nbytes or min(self.nbytes, other.nbytes),
                                     _stream_ptr(stream)))

    def download(self, dtype=np.uint8) -> np.ndarray:
        """Synchronous copy of the whole buffer to the host (as one image row
        of `nbytes` bytes)."""
        out = np.empty(self.nbytes, np.uint8)
        row = _Image(self.ptr, self.nbytes, 1, 1)
        _check(lib.ssim2_image_download(ctypes.byref(row), out.ctypes.data, self.nbytes))
        return out.view(dtype)

    def free(self):
        if self.ptr:
            lib.ssim2_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def srgb_to_linear(src: DeviceImage, dst: DeviceImage, stream: Stream | None = None):
    """Kernel::srgb_to_linear (crates/ssimulacra2-cuda/src/kernel.rs:36-57)."""
    _check(lib.ssim2_srgb_to_linear(src.as_struct(), dst.as_struct(), _stream_ptr(stream)))


def downscale_by_2(src: DeviceImage, dst: DeviceImage, stream: Stream | None = None):
    """Kernel::downscale_by_2 (crates/ssimulacra2-cuda/src/kernel.rs:59-83): f32 images, dst half size (rounded up)."""
    _check(lib.ssim2_downscale_by_2(src.as_struct(), dst.as_struct(), _stream_ptr(stream)))


def linear_to_xyb(src: DeviceImage, dst: DeviceImage, stream: Stream | None = None):
    """Kernel::linear_to_xyb (crates/ssimulacra2-cuda/src/kernel.rs:161-184): f32 linear RGB to positive XYB."""
    _check(lib.ssim2_linear_to_xyb(src.as_struct(), dst.as_struct(), _stream_ptr(stream)))


class Kernel:
    """Ssimulacra2::kernel() (crates/ssimulacra2-cuda/src/lib.rs:199-204): the
    individual stage kernels, as module functions."""

    srgb_to_linear = staticmethod(lambda src, dst, stream=None: srgb_to_linear(src, dst, stream))
    downscale_by_2 = staticmethod(lambda src, dst, stream=None: downscale_by_2(src, dst, stream))
    linear_to_xyb = staticmethod(lambda src, dst, stream=None: linear_to_xyb(src, dst, stream))


class Ssimulacra2:
    """ssimulacra2_cuda::Ssimulacra2 for MI355X (crates/ssimulacra2-cuda/src/lib.rs:27-602).

    Construct with the frame size (`Ssimulacra2(width, height)`) or, as in the
    reference, with the two linear images (`Ssimulacra2.new(ref_lin, dis_lin)`),
    which are then the default inputs of `compute` / `compute_sync`.
    """

    def __init__(self, width: int, height: int, memory_budget: int = 0):
        opts = _Options(ctypes.sizeof(_Options), 0, memory_budget)
        h = ctypes.c_void_p()
        _check(lib.ssim2_create(width, height, ctypes.byref(opts), ctypes.byref(h)))
        self._ctx = h
        self.width, self.height = width, height
        self._bound = None

    # -- construction / lifecycle (lib.rs:80-197, 1195-1202)
    @classmethod
    def new(cls, ref_linear: DeviceImage, dis_linear: DeviceImage, stream=None, memory_budget: int = 0):
        if (ref_linear.width, ref_linear.height) != (dis_linear.width, dis_linear.height):
            raise Ssim2Error(2, "reference and distorted sizes differ")
        s = cls(ref_linear.width, ref_linear.height, memory_budget)
        s._bound = (ref_linear, dis_linear)
        return s

    def close(self):
        if getattr(self, "_ctx", None):
            lib.ssim2_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def mem_usage(self) -> int:
        v = ctypes.c_size_t()
        _check(lib.ssim2_mem_usage(self._ctx, ctypes.byref(v)))
        return v.value

    # -- scoring (lib.rs:329-429)
    def compute(self, stream=None, ref_linear=None, dis_linear=None):
        ref_linear, dis_linear = self._inputs(ref_linear, dis_linear)
        _check(lib.ssim2_compute_linear_async(self._ctx, ref_linear.as_struct(), dis_linear.as_struct(),
                                              _stream_ptr(stream)))

    def get_score(self) -> float:
        v = ctypes.c_double()
        _check(lib.ssim2_get_score(self._ctx, ctypes.byref(v)))
        return v.value

    def compute_sync(self, stream=None, ref_linear=None, dis_linear=None) -> float:
        ref_linear, dis_linear = self._inputs(ref_linear, dis_linear)
        v = ctypes.c_double()
        _check(lib.ssim2_compute_linear_sync(self._ctx, ref_linear.as_struct(), dis_linear.as_struct(),
                                             _stream_ptr(stream), ctypes.byref(v)))
        return v.value

    def compute_srgb_sync(self, src_ref: DeviceImage, src_dis: DeviceImage, ref_linear=None,
                          dis_linear=None, stream=None) -> float:
        v = ctypes.c_double()
        _check(lib.ssim2_compute_srgb_sync(
            self._ctx, src_ref.as_struct(), src_dis.as_struct(),
            ref_linear.as_struct() if ref_linear is not None else None,
            dis_linear.as_struct() if dis_linear is not None else None,
            _stream_ptr(stream), ctypes.byref(v)))
        return v.value

    def compute_srgb(self, src_ref: DeviceImage, src_dis: DeviceImage, stream=None):
        """Asynchronous u8 path; pair with get_score()."""
        _check(lib.ssim2_compute_srgb_async(self._ctx, src_ref.as_struct(), src_dis.as_struct(),
                                            _stream_ptr(stream)))

    def compute_yuv420(self, ref: DeviceYuv, dis: DeviceYuv, stream=None):
        """Async score of a YUV 4:2:0 frame pair (conversion fused into the first kernel)."""
        _check(lib.ssim2_compute_yuv420_async(self._ctx, ref.as_struct(), dis.as_struct(), _stream_ptr(stream)))

    def compute_yuv420_sync(self, ref: DeviceYuv, dis: DeviceYuv, stream=None) -> float:
        out = ctypes.c_double()
        _check(lib.ssim2_compute_yuv420_sync(self._ctx, ref.as_struct(), dis.as_struct(), _stream_ptr(stream),
                                             ctypes.byref(out)))
        return out.value

    def compute_from_cpu_srgb_sync(self, ref_bytes: np.ndarray, dis_bytes: np.ndarray, stream=None) -> float:
        r = np.ascontiguousarray(ref_bytes, dtype=np.uint8)
        d = np.ascontiguousarray(dis_bytes, dtype=np.uint8)
        row = self.width * 3
        if r.size != row * self.height or d.size != row * self.height:
            raise Ssim2Error(2, "host buffers do not match the context size")
        v = ctypes.c_double()
        _check(lib.ssim2_compute_from_host_srgb(self._ctx, r.ctypes.data, d.ctypes.data, row,
                                                _stream_ptr(stream), ctypes.byref(v)))
        return v.value

    # -- cached reference (lib.rs:440-602)
    def prepare_reference_from_cpu_srgb(self, ref_bytes: np.ndarray, stream=None):
        """Upload host sRGB bytes (h, w, 3) and cache them as the reference (lib.rs:373-388)."""
        a = np.ascontiguousarray(ref_bytes, np.uint8)
        _check(lib.ssim2_prepare_reference_from_host_srgb(self._ctx, a.ctypes.data, a.strides[0],
                                                          _stream_ptr(stream)))

    def compute_from_cpu_dis_only_srgb_sync(self, dis_bytes: np.ndarray, stream=None) -> float:
        """Score host sRGB bytes against the cached reference (lib.rs:347-371)."""
        a = np.ascontiguousarray(dis_bytes, np.uint8)
        v = ctypes.c_double()
        _check(lib.ssim2_compute_from_host_dis_srgb(self._ctx, a.ctypes.data, a.strides[0], _stream_ptr(stream),
                                                    ctypes.byref(v)))
        return v.value

    def set_reference_linear(self, ref_linear: DeviceImage, stream=None):
        _check(lib.ssim2_set_reference_linear(self._ctx, ref_linear.as_struct(), _stream_ptr(stream)))

    def set_reference_srgb(self, ref_srgb: DeviceImage, stream=None):
        _check(lib.ssim2_set_reference_srgb(self._ctx, ref_srgb.as_struct(), _stream_ptr(stream)))

    def compute_with_reference_linear(self, ref_linear, dis_linear: DeviceImage, stream=None) -> float:
        v = ctypes.c_double()
        _check(lib.ssim2_compute_with_reference_linear(
            self._ctx, ref_linear.as_struct() if ref_linear is not None else None,
            dis_linear.as_struct(), _stream_ptr(stream), ctypes.byref(v)))
        return v.value

    def compute_with_reference_srgb(self, dis_srgb: DeviceImage, stream=None) -> float:
        v = ctypes.c_double()
        _check(lib.ssim2_compute_with_reference_srgb(self._ctx, dis_srgb.as_struct(), _stream_ptr(stream),
                                                     ctypes.byref(v)))
        return v.value

    def clear_reference(self):
        _check(lib.ssim2_clear_reference(self._ctx))

    def has_reference(self) -> bool:
        v = ctypes.c_int()
        _check(lib.ssim2_has_reference(self._ctx, ctypes.byref(v)))
        return bool(v.value)

    def kernel(self) -> "Kernel":
        """Ssimulacra2::kernel (lib.rs:199-204)."""
        return Kernel()

    # -- diagnostics
    def time_kernels(self, src_ref: DeviceImage, src_dis: DeviceImage, reps: int = 10, stream=None):
        """Mean ms per launch of {pyramid, rows, cols, finalize} (HIP events on `stream`)."""
        ms = (ctypes.c_float * 4)()
        _check(lib.ssim2_time_kernels(self._ctx, src_ref.as_struct(), src_dis.as_struct(),
                                      _stream_ptr(stream), reps, ms))
        return dict(zip(("pyramid", "rows", "cols", "finalize"), [float(v) for v in ms]))

    def plane(self, kind: str, scale: int, index: int) -> np.ndarray:
        """An intermediate f32 plane of the last computation: kind "xyb"
        (index 0..5: reference X, Y, B, distorted X, Y, B) or "rows" (index
        channel * 5 + quantity: ref^2, dis^2, ref*dis, ref, dis)."""
        w, h = self.width, self.height
        for _ in range(scale):
            w, h = (w + 1) // 2, (h + 1) // 2
        out = np.empty((h, w), np.float32)
        _check(lib.ssim2_get_plane(self._ctx, {"xyb": 0, "rows": 1}[kind], scale, index, out.ctypes.data, w))
        return out

    def sums(self):
        """(sums[6,3,6], dims[6,2], nscales) behind the last score."""
        s = (ctypes.c_double * 108)()
        d = (ctypes.c_int * 12)()
        n = ctypes.c_int()
        _check(lib.ssim2_get_sums(self._ctx, s, d, ctypes.byref(n)))
        return (np.array(s[:], dtype=np.float64).reshape(6, 3, 6),
                np.array(d[:], dtype=np.int32).reshape(6, 2), n.value)

    def _inputs(self, ref_linear, dis_linear):
        if ref_linear is None or dis_linear is None:
            if self._bound is None:
                raise Ssim2Error(8, "no images bound: pass ref_linear/dis_linear or use Ssimulacra2.new")
            return self._bound
        return ref_linear, dis_linear


class Ssim2Batch:
    """Batched fast form: many sRGB pairs of any sizes per call (include/ssim2hip.h ssim2_batch_*).

    Scores follow the FP64 evaluation of the metric (exact FIR blur, shifted
    statistics); they are not bit-identical to `Ssimulacra2`'s exact path.
    """

    def __init__(self, max_pairs: int, max_pixels: int, storage: str = "f32", memory_budget: int = 0):
        o = _BatchOptions(ctypes.sizeof(_BatchOptions), STORAGE[storage], max_pairs, 0, max_pixels, memory_budget)
        h = ctypes.c_void_p()
        _check(lib.ssim2_batch_create(ctypes.byref(o), ctypes.byref(h)))
        self._b = h
        self.storage = storage
        self.max_pairs = max_pairs
        self._last = None

    def close(self):
        if getattr(self, "_b", None):
            lib.ssim2_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @staticmethod
    def _pairs(pairs):
        arr = (_Pair * len(pairs))()
        for i, (r, d) in enumerate(pairs):
            arr[i].ref = r._img
            arr[i].dis = d._img
        return arr

    def compute_async(self, pairs, stream=None):
        """pairs: sequence of (ref DeviceImage u8, dis DeviceImage u8)."""
        arr = self._pairs(pairs)
        self._last = (arr, len(pairs))
        _check(lib.ssim2_batch_compute_async(self._b, arr, len(pairs), _stream_ptr(stream)))

    def add_reference(self, ref, stream=None) -> int:
        """Compute and keep the pyramid of a reference DeviceImage (u8); the
        id names it in compute_ref pairs (crates/ssimulacra2-cuda/src/lib.rs
        set_reference, batched)."""
        v = ctypes.c_uint32()
        _check(lib.ssim2_batch_add_reference(self._b, ctypes.byref(ref._img), _stream_ptr(stream), ctypes.byref(v)))
        return v.value

    def remove_reference(self, ref_id: int):
        _check(lib.ssim2_batch_remove_reference(self._b, ref_id))

    def compute_ref_async(self, pairs, stream=None):
        """pairs: sequence of (reference id, dis DeviceImage u8)."""
        arr = (_RefPair * len(pairs))()
        for i, (rid, d) in enumerate(pairs):
            arr[i].ref_id = rid
            arr[i].dis = d._img
        self._last = (arr, len(pairs))
        _check(lib.ssim2_batch_compute_ref_async(self._b, arr, len(pairs), _stream_ptr(stream)))

    def compute_ref(self, pairs, stream=None):
        self.compute_ref_async(pairs, stream)
        return self.get_scores()

    def compute_yuv_async(self, pairs, stream=None):
        """pairs: sequence of (ref DeviceYuv, dis DeviceYuv)."""
        arr = (_YuvPair * len(pairs))()
        for i, (r, d) in enumerate(pairs):
            arr[i].ref = r._img
            arr[i].dis = d._img
        self._last = (arr, len(pairs))
        _check(lib.ssim2_batch_compute_yuv_async(self._b, arr, len(pairs), _stream_ptr(stream)))

    def compute_yuv(self, pairs, stream=None):
        self.compute_yuv_async(pairs, stream)
        return self.get_scores()

    def get_scores(self):
        n = self._last[1]
        out = (ctypes.c_double * n)()
        _check(lib.ssim2_batch_get_scores(self._b, out, n))
        return np.array(out[:], dtype=np.float64)

    def compute(self, pairs, stream=None):
        self.compute_async(pairs, stream)
        return self.get_scores()

    def done(self) -> bool:
        v = ctypes.c_int()
        _check(lib.ssim2_batch_query(self._b, ctypes.byref(v)))
        return bool(v.value)

    def sums(self, index: int):
        s = (ctypes.c_double * 108)()
        d = (ctypes.c_int * 12)()
        _check(lib.ssim2_batch_get_sums(self._b, index, s, d))
        return np.array(s[:]).reshape(6, 3, 6), np.array(d[:], dtype=np.int32).reshape(6, 2)

    def set_tiling(self, tile_cfg: int, grid_mult: int):
        _check(lib.ssim2_batch_set_tiling(self._b, tile_cfg, grid_mult))

    def tiling(self):
        a, b = ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib.ssim2_batch_get_tiling(self._b, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def autotune(self, pairs, reps: int = 3, stream=None) -> str:
        arr = self._pairs(pairs)
        buf = ctypes.create_string_buffer(4096)
        _check(lib.ssim2_batch_autotune(self._b, arr, len(pairs), _stream_ptr(stream), reps, buf, 4096))
        return buf.value.decode()

    def mem_usage(self) -> int:
        v = ctypes.c_size_t()
        _check(lib.ssim2_batch_mem_usage(self._b, ctypes.byref(v)))
        return v.value


class Ssim2Server:
    """Continuous-batching scorer (include/ssim2hip.h ssim2_server_*).

    submit() is thread-safe and returns at once; results come back through
    poll() with host timestamps. Submitted images must stay alive until their
    result has been polled (this wrapper keeps references until then).
    """

    def __init__(self, storage="f16", slots=3, max_pairs_per_batch=256, max_pixels_per_batch=64 << 20,
                 hbm_budget=64 << 30, tile_cfg=0, grid_mult=0):
        o = _ServerOptions(ctypes.sizeof(_ServerOptions), STORAGE[storage], slots, max_pairs_per_batch,
                           max_pixels_per_batch, hbm_budget, tile_cfg, grid_mult)
        h = ctypes.c_void_p()
        _check(lib.ssim2_server_create(ctypes.byref(o), ctypes.byref(h)))
        self._s = h
        self._keep = {}
        self._buf = (_Result * 4096)()

    def submit(self, ref: DeviceImage, dis: DeviceImage, request_id: int):
        p = _Pair(ref._img, dis._img)
        _check(lib.ssim2_server_submit(self._s, ctypes.byref(p), request_id))
        self._keep[request_id] = (ref, dis)

    def add_reference(self, ref: DeviceImage) -> int:
        """Keep the pyramid of a reference resident for every slot."""
        v = ctypes.c_uint32()
        _check(lib.ssim2_server_add_reference(self._s, ctypes.byref(ref._img), ctypes.byref(v)))
        return v.value

    def remove_reference(self, ref_id: int):
        _check(lib.ssim2_server_remove_reference(self._s, ref_id))

    def submit_ref(self, ref_id: int, dis: DeviceImage, request_id: int):
        _check(lib.ssim2_server_submit_ref(self._s, ref_id, ctypes.byref(dis._img), request_id))
        self._keep[request_id] = (dis,)

    def poll(self):
        """Finished results as a list of (id, score, t_submit, t_done, status)."""
        out = []
        while True:
            n = ctypes.c_uint32()
            _check(lib.ssim2_server_poll(self._s, self._buf, len(self._buf), ctypes.byref(n)))
            for i in range(n.value):
                r = self._buf[i]
                out.append((r.request_id, r.score, r.submit_time_s, r.done_time_s, STATUS.get(r.status, r.status)))
                self._keep.pop(r.request_id, None)
            if n.value < len(self._buf):
                return out

    def drain(self, timeout_s: float = 0.0):
        _check(lib.ssim2_server_drain(self._s, timeout_s))

    def stats(self) -> dict:
        st = _ServerStats()
        _check(lib.ssim2_server_stats(self._s, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in _ServerStats._fields_ if not k.startswith(("struct", "reserved"))}

    def close(self):
        if getattr(self, "_s", None):
            lib.ssim2_server_destroy(self._s)
            self._s = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
