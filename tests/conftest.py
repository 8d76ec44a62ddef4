"""Shared fixtures: the package path, the CPU oracle (test infrastructure) and
the synthetic generators."""
import contextlib
import ctypes
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG_DIR = ROOT / "turbo-metrics_amd"
ORACLE_SO = ROOT / "oracle" / "libssim2_oracle.so"
GOLDEN = Path(__file__).resolve().parent / "golden"
if str(PKG_DIR) not in sys.path:
    sys.path.insert(0, str(PKG_DIR))
if str(ROOT / "oracle") not in sys.path:
    sys.path.insert(0, str(ROOT / "oracle"))

MIRROR, F64, F32_FIR = 0, 1, 2


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


class Oracle:
    """ctypes view of oracle/libssim2_oracle.so (test infrastructure only)."""

    def __init__(self, path=ORACLE_SO):
        if not path.exists():
            raise RuntimeError(f"{path} missing: run __graft_entry__.build()")
        L = ctypes.CDLL(str(path))
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.oracle_score_srgb.argtypes = [vp, vp, i, i, i, ctypes.POINTER(d), vp, vp]
        L.oracle_score_linear.argtypes = [vp, vp, i, i, i, ctypes.POINTER(d), vp, vp]
        L.oracle_score_from_sums.argtypes = [vp, vp, i]
        L.oracle_score_from_sums.restype = d
        L.oracle_scale_dims.argtypes = [i, i, vp]
        L.oracle_srgb_lut.argtypes = [vp]
        L.oracle_weights.argtypes = [vp]
        L.oracle_linear_to_xyb_f32.argtypes = [vp, ctypes.c_size_t, vp]
        L.oracle_linear_to_xyb_f64.argtypes = [vp, ctypes.c_size_t, vp]
        L.oracle_downscale_f32.argtypes = [vp, i, i, vp]
        for n in ("oracle_blur_f32", "oracle_blur_rows_f32", "oracle_blur_cols_f32", "oracle_blur_fir_f32",
                  "oracle_blur_f64"):
            getattr(L, n).argtypes = [vp, i, i, vp]
        L.oracle_map_sums_f32.argtypes = [i] + [vp] * 7 + [vp]
        L.oracle_map_sums_fast.argtypes = [i] + [vp] * 7 + [vp]
        L.oracle_round_f16.argtypes = [vp, ctypes.c_size_t, vp]
        L.oracle_plant.argtypes = [i]
        L.oracle_score_srgb_many.argtypes = [vp, vp, vp, vp, i, i, i, vp]
        L.oracle_bt709_coeffs.argtypes = [vp]
        L.oracle_bt709_eotf.argtypes = [d]
        L.oracle_bt709_eotf.restype = d
        L.oracle_yuv420p10_to_linear_f64.argtypes = [vp, vp, vp, i, i, vp]
        L.oracle_score_linear_f64.argtypes = [vp, vp, i, i, ctypes.POINTER(d)]
        L.oracle_score_yuv420p10_f64.argtypes = [vp] * 6 + [i, i, ctypes.POINTER(d)]
        self.L = L

    def score_srgb(self, ref, dis, mode=MIRROR, with_sums=False):
        h, w = ref.shape[:2]
        r = np.ascontiguousarray(ref, np.uint8)
        q = np.ascontiguousarray(dis, np.uint8)
        s = ctypes.c_double()
        sums = np.zeros(108)
        dims = np.zeros(12, np.int32)
        n = self.L.oracle_score_srgb(r.ctypes.data, q.ctypes.data, w, h, mode, ctypes.byref(s),
                                     sums.ctypes.data, dims.ctypes.data)
        if with_sums:
            return s.value, sums.reshape(6, 3, 6), dims.reshape(6, 2), n
        return s.value

    def score_linear(self, ref, dis, mode=MIRROR, with_sums=False):
        h, w = ref.shape[:2]
        r = np.ascontiguousarray(ref, np.float32)
        q = np.ascontiguousarray(dis, np.float32)
        s = ctypes.c_double()
        sums = np.zeros(108)
        dims = np.zeros(12, np.int32)
        n = self.L.oracle_score_linear(r.ctypes.data, q.ctypes.data, w, h, mode, ctypes.byref(s),
                                       sums.ctypes.data, dims.ctypes.data)
        if with_sums:
            return s.value, sums.reshape(6, 3, 6), dims.reshape(6, 2), n
        return s.value

    def lut(self):
        out = np.zeros(256, np.float32)
        self.L.oracle_srgb_lut(out.ctypes.data)
        return out

    def weights(self):
        out = np.zeros(108, np.float64)
        self.L.oracle_weights(out.ctypes.data)
        return out

    def plane_op(self, name, plane):
        a = np.ascontiguousarray(plane)
        h, w = a.shape
        out = np.zeros_like(a)
        getattr(self.L, name)(a.ctypes.data, w, h, out.ctypes.data)
        return out

    def xyb_f32(self, rgb):
        a = np.ascontiguousarray(rgb, np.float32)
        out = np.zeros_like(a)
        self.L.oracle_linear_to_xyb_f32(a.ctypes.data, a.size // 3, out.ctypes.data)
        return out

    def xyb_f64(self, rgb):
        a = np.ascontiguousarray(rgb, np.float64)
        out = np.zeros_like(a)
        self.L.oracle_linear_to_xyb_f64(a.ctypes.data, a.size // 3, out.ctypes.data)
        return out

    def downscale(self, rgb):
        a = np.ascontiguousarray(rgb, np.float32)
        h, w = a.shape[:2]
        out = np.zeros(((h + 1) // 2, (w + 1) // 2, 3), np.float32)
        self.L.oracle_downscale_f32(a.ctypes.data, w, h, out.ctypes.data)
        return out

    def bt709(self):
        """(y, r, b, g1, g2, Kr, Kb) of the double derivation."""
        c = np.zeros(7)
        self.L.oracle_bt709_coeffs(c.ctypes.data)
        return c

    def yuv_to_linear(self, y, cb, cr):
        """10-bit planes -> (h, w, 3) float64 linear RGB."""
        y = np.ascontiguousarray(y, np.uint16)
        cb = np.ascontiguousarray(cb, np.uint16)
        cr = np.ascontiguousarray(cr, np.uint16)
        h, w = y.shape
        out = np.empty((h, w, 3))
        self.L.oracle_yuv420p10_to_linear_f64(y.ctypes.data, cb.ctypes.data, cr.ctypes.data, w, h, out.ctypes.data)
        return out

    def score_linear_f64(self, ref, dis):
        a = np.ascontiguousarray(ref, np.float64)
        b = np.ascontiguousarray(dis, np.float64)
        h, w = a.shape[:2]
        s = ctypes.c_double()
        self.L.oracle_score_linear_f64(a.ctypes.data, b.ctypes.data, w, h, ctypes.byref(s))
        return s.value

    def score_yuv_f64(self, ref, dis):
        """ref, dis: (Y, Cb, Cr) 10-bit planes."""
        planes = [np.ascontiguousarray(p, np.uint16) for p in (*ref, *dis)]
        h, w = planes[0].shape
        s = ctypes.c_double()
        self.L.oracle_score_yuv420p10_f64(*[p.ctypes.data for p in planes], w, h, ctypes.byref(s))
        return s.value

    def map_sums(self, src, dis, mu1, mu2, s11, s22, s12, fast=False):
        """Error-map sums {ssim S1, S4, art S1, S4, det S1, S4} of one channel:
        the mirror's (cpu.rs:581-683) or, with fast=True, the fast form's f32 maps."""
        arrs = [np.ascontiguousarray(a, np.float32) for a in (src, dis, mu1, mu2, s11, s22, s12)]
        out = np.zeros(6)
        fn = self.L.oracle_map_sums_fast if fast else self.L.oracle_map_sums_f32
        fn(arrs[0].size, *[a.ctypes.data for a in arrs], out.ctypes.data)
        return out

    def round_f16(self, x):
        a = np.ascontiguousarray(x, np.float32)
        out = np.empty_like(a)
        self.L.oracle_round_f16(a.ctypes.data, a.size, out.ctypes.data)
        return out

    @contextlib.contextmanager
    def planted(self, which):
        """Run the oracle with planted defect `which` (see ssim2_oracle.cpp g_plant)."""
        self.L.oracle_plant(which)
        try:
            yield
        finally:
            self.L.oracle_plant(0)

    def scale_dims(self, w, h):
        d = np.zeros(12, np.int32)
        n = self.L.oracle_scale_dims(w, h, d.ctypes.data)
        return [tuple(d[2 * i:2 * i + 2]) for i in range(n)]


@pytest.fixture(scope="session")
def oracle():
    return Oracle()


@pytest.fixture(scope="session")
def synth():
    from ssimulacra2_hip import synth as s
    return s
