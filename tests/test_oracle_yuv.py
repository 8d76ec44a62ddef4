"""CPU pins of the oracle's YUV 4:2:0 -> linear conversion (the video input,
crates/cuda-colorspace-kernel/src/biplanar.rs:8-70 and lib.rs:185-246).

The reference holds no test vectors for this path, so the conversion is pinned
by what the mathematics and the standard fix: BT.709's published Kr/Kb, grey for
neutral chroma, black and white at the limited-range floor and ceiling, the
inverse OETF's continuity, and a forward encode -> decode round trip (a swapped
Cb/Cr, a wrong sign or a dropped chroma term fails it). The HIP kernels' f32
coefficients are re-derived here in float32 by the reference's own procedure
(primaries -> Kr/Kb -> matrix, lib.rs:185-201, const_algebra.rs:21-31)."""
import re

import numpy as np
import pytest

from conftest import ROOT


def test_kr_kb_are_bt709(oracle):
    c = oracle.bt709()
    # Rec. ITU-R BT.709-6 item 3.2: Kr = 0.2126, Kb = 0.0722 (rounded to 4 places)
    assert abs(c[5] - 0.2126) < 5e-5
    assert abs(c[6] - 0.0722) < 5e-5


def test_neutral_chroma_is_grey_and_range_ends(oracle):
    ys = np.arange(0, 1024, dtype=np.uint16)
    y = np.tile(ys, (2, 1))                         # 2 x 1024 frame
    cb = np.full((1, 512), 512, np.uint16)
    cr = np.full((1, 512), 512, np.uint16)
    lin = oracle.yuv_to_linear(y, cb, cr)
    assert np.array_equal(lin[..., 0], lin[..., 1]) and np.array_equal(lin[..., 0], lin[..., 2])
    row = lin[0, :, 0]
    assert np.all(row[:65] == 0.0)                  # at or below the floor (64)
    assert abs(row[940] - 1.0) < 1e-12              # the ceiling (940) is white
    assert np.all(row[940:] == 1.0)                 # clamped above
    assert np.all(np.diff(row[64:941]) > 0)         # strictly increasing in between


def test_eotf_continuous_and_inverts_oetf(oracle):
    beta = 0.018053968510807
    alpha = 1 + 5.5 * beta
    thr = alpha * beta ** 0.45 - (alpha - 1)
    assert abs(thr - 0.08124285829863521) < 1e-12   # the threshold the reference states
    assert abs(oracle.L.oracle_bt709_eotf(thr) - thr / 4.5) < 1e-12
    for L in np.linspace(0.0, 1.0, 101):
        v = alpha * L ** 0.45 - (alpha - 1) if L >= beta else 4.5 * L
        assert abs(oracle.L.oracle_bt709_eotf(v) - L) < 1e-12


def _encode_10bit(rgb, kr, kb):
    """Linear RGB (h, w, 3), constant over each 2x2 quad -> 10-bit limited Y, Cb, Cr."""
    beta = 0.018053968510807
    alpha = 1 + 5.5 * beta
    e = np.where(rgb >= beta, alpha * np.power(rgb, 0.45) - (alpha - 1), 4.5 * rgb)
    kg = 1 - kr - kb
    yp = kr * e[..., 0] + kg * e[..., 1] + kb * e[..., 2]
    pb = (e[..., 2] - yp) / (2 * (1 - kb))
    pr = (e[..., 0] - yp) / (2 * (1 - kr))
    Y = np.round(64 + 876 * yp).astype(np.uint16)
    Cb = np.round(512 + 896 * pb[::2, ::2]).astype(np.uint16)
    Cr = np.round(512 + 896 * pr[::2, ::2]).astype(np.uint16)
    return Y, Cb, Cr


def test_round_trip(oracle):
    rng = np.random.default_rng(3)
    quads = rng.uniform(0.02, 0.98, (16, 24, 3))
    rgb = np.repeat(np.repeat(quads, 2, axis=0), 2, axis=1)   # 32 x 48
    c = oracle.bt709()
    Y, Cb, Cr = _encode_10bit(rgb, c[5], c[6])
    back = oracle.yuv_to_linear(Y, Cb, Cr)
    # 10-bit quantisation of the gamma-encoded values: half a code is
    # <= 0.5/876 in E', i.e. < 3e-3 in linear light over [0.02, 0.98]
    assert np.max(np.abs(back - rgb)) < 3e-3
    # swapping the chroma planes must be caught
    assert np.max(np.abs(oracle.yuv_to_linear(Y, Cr, Cb) - rgb)) > 0.05


def test_odd_size_uses_the_quads_chroma(oracle):
    rng = np.random.default_rng(5)
    w, h = 7, 5
    Y = rng.integers(64, 941, (h, w)).astype(np.uint16)
    Cb = rng.integers(64, 961, (3, 4)).astype(np.uint16)
    Cr = rng.integers(64, 961, (3, 4)).astype(np.uint16)
    full = oracle.yuv_to_linear(Y, Cb, Cr)
    for py in range(h):
        for px in range(w):
            one = oracle.yuv_to_linear(np.full((2, 2), Y[py, px], np.uint16),
                                       Cb[py // 2:py // 2 + 1, px // 2:px // 2 + 1],
                                       Cr[py // 2:py // 2 + 1, px // 2:px // 2 + 1])
            assert np.array_equal(full[py, px], one[0, 0])


def _kernel_constants():
    text = (ROOT / "turbo-metrics_amd" / "csrc" / "ssim2_math.h").read_text()
    out = {}
    for name in ("kYuvY", "kYuvR", "kYuvB", "kYuvG1", "kYuvG2", "kBt709AlphaM1", "kBt709Alpha",
                 "kBt709Threshold", "kBt709InvGamma"):
        m = re.search(name + r"\s*=\s*(-?0x[0-9a-fA-F.]+p[-+]?\d+)f", text)
        assert m, name
        out[name] = float.fromhex(m.group(1))
    return out


def test_kernel_f32_coefficients_follow_the_reference_derivation(oracle):
    f = np.float32

    def xyz(x, y):
        x, y = f(x), f(y)
        return (x / y, f(1.0), (f(1.0) - x - y) / y)

    def cross(a, b):
        return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])

    def dot(a, b):
        return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]

    r, g, b, w = xyz(0.64, 0.33), xyz(0.30, 0.60), xyz(0.15, 0.06), xyz(0.3127, 0.3290)
    mul = f(1.0) / dot((r[0], g[0], b[0]), cross((r[1], g[1], b[1]), (r[2], g[2], b[2])))
    kr, kb = dot(w, cross(g, b)) * mul, dot(w, cross(r, g)) * mul
    kg = f(1.0) - kr - kb
    cr = f((240 - 16) << 8)
    want = {
        "kYuvY": f(1.0) / f((235 - 16) << 8),
        "kYuvR": f(2.0) * (f(1.0) - kr) * f(1.0) / cr,
        "kYuvB": f(2.0) * (f(1.0) - kb) * f(1.0) / cr,
        "kYuvG1": -f(2.0) * (f(1.0) - kb) * kb / kg * f(1.0) / cr,
        "kYuvG2": -f(2.0) * (f(1.0) - kr) * kr / kg * f(1.0) / cr,
    }
    beta = f(0.018053968510807)
    alpha = f(1.0) + f(5.5) * beta
    want.update(kBt709AlphaM1=alpha - f(1.0), kBt709Alpha=alpha, kBt709Threshold=f(0.08124285829863521),
                kBt709InvGamma=f(1.0) / f(0.45))
    got = _kernel_constants()
    for k, v in want.items():
        assert got[k] == float(v), (k, got[k], float(v))
    # and within f32 rounding of the oracle's double derivation
    c = oracle.bt709()
    for k, dv in zip(("kYuvY", "kYuvR", "kYuvB", "kYuvG1", "kYuvG2"), c[:5]):
        assert abs(got[k] - dv) <= 4e-7 * abs(dv), k


def test_identical_frames_score_100(oracle, synth):
    ref, _ = synth.yuv420_pair(11, 64, 48, 0)
    assert oracle.score_yuv_f64(ref, ref) == 100.0


def test_banding_lowers_the_score(oracle, synth):
    ref, dis = synth.yuv420_pair(11, 128, 96, 0)
    s = oracle.score_yuv_f64(ref, dis)
    assert 40.0 < s < 99.0
    assert oracle.score_linear_f64(oracle.yuv_to_linear(*ref), oracle.yuv_to_linear(*dis)) == s
