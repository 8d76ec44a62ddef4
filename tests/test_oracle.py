"""Pins for the CPU oracle (oracle/ssim2_oracle.cpp) — CPU only.

Each check is chosen so that a plausible slip in the restatement of
crates/ssimulacra2-cuda/examples/cpu.rs fails it: a wrong table entry, a
transposed recurrence form, a dropped FIR tap, a wrong scale count, a
mis-ordered weight. None re-types the oracle's formulas.
"""
import json
from pathlib import Path
import math
import struct

import numpy as np
import pytest

from conftest import F32_FIR, F64, GOLDEN, MIRROR, ROOT

TABLES = json.loads((GOLDEN / "reference_tables.json").read_text())


def f32_bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


# ---------------------------------------------------------------- tables


def test_lut_equals_reference_table(oracle):
    """sRGB LUT: bit-equal to cpu.rs:20-277 (and srgb.rs:18-51, checked at extraction)."""
    np.testing.assert_array_equal(f32_bits(oracle.lut()), np.array(TABLES["srgb8_lut_f32_bits"], np.uint32))


def test_lut_is_close_to_srgb_eotf(oracle):
    """The table is the sRGB EOTF (threshold 0.04045, gamma 2.4) to < 4e-7 (a closed form)."""
    x = np.arange(256) / 255.0
    eotf = np.where(x <= 0.04045, x / 12.92, ((x + 0.055) / 1.055) ** 2.4)
    assert np.abs(oracle.lut() - eotf).max() < 4e-7
    assert oracle.lut()[0] == 0.0 and oracle.lut()[255] == 1.0


def test_weights_equal_reference_table(oracle):
    got = oracle.weights().view(np.uint64)
    np.testing.assert_array_equal(got, np.array(TABLES["weight_f64_bits"], np.uint64))


def _parse_hex_table(path, macro):
    text = path.read_text()
    i = text.index(f"#define {macro}")
    j = text.index("}", i)
    body = text[i:j].split("{", 1)[1]
    return [float.fromhex(t.strip().rstrip("f")) for t in body.replace("\\", " ").split(",") if t.strip()]


def test_product_tables_equal_reference():
    """The HIP library's copy of the tables (csrc/ssim2_tables.h) is the same data."""
    hdr = ROOT / "turbo-metrics_amd" / "csrc" / "ssim2_tables.h"
    lut = np.array(_parse_hex_table(hdr, "SSIM2_SRGB8_LUT_INIT"), np.float32)
    w = np.array(_parse_hex_table(hdr, "SSIM2_WEIGHTS_INIT"), np.float64)
    np.testing.assert_array_equal(f32_bits(lut), np.array(TABLES["srgb8_lut_f32_bits"], np.uint32))
    np.testing.assert_array_equal(w.view(np.uint64), np.array(TABLES["weight_f64_bits"], np.uint64))


def test_iir_constants_rederived():
    """oracle/derive_iir.py solves Charalampidis (2016) and lands on cpu.rs:931-948."""
    import derive_iir

    c = derive_iir.f32_constants()
    assert c["RADIUS"] == TABLES["iir_radius"] == 5
    for name, bits in TABLES["iir_f32_bits"].items():
        assert struct.unpack("<I", struct.pack("<f", float(c[name])))[0] == bits, name


def test_kernel_constants_match_reference():
    """ssim2_math.h carries the same f32 recurrence constants."""
    text = (ROOT / "turbo-metrics_amd" / "csrc" / "ssim2_math.h").read_text()
    want = {k: struct.unpack("<f", struct.pack("<I", v))[0] for k, v in TABLES["iir_f32_bits"].items()}
    pairs = {"kMulIn1": "MUL_IN_1", "kMulIn3": "MUL_IN_3", "kMulIn5": "MUL_IN_5",
             "kMulPrev1": "MUL_PREV_1", "kMulPrev3": "MUL_PREV_3", "kMulPrev5": "MUL_PREV_5",
             "kVertMulPrev1": "VERT_MUL_PREV_1", "kVertMulPrev3": "VERT_MUL_PREV_3",
             "kVertMulPrev5": "VERT_MUL_PREV_5"}
    for cname, rname in pairs.items():
        tok = text.split(cname + " = ", 1)[1].split(",")[0].split(";")[0].strip().rstrip("f")
        assert np.float32(float.fromhex(tok)) == np.float32(want[rname]), cname


# ---------------------------------------------------------------- blur


def _fir_brute(plane, taps):
    """Zero-padded separable 9-tap convolution in double, by brute force."""
    h, w = plane.shape
    k = np.concatenate([taps[:0:-1], taps])  # [h4..h1, h0, h1..h4]
    pad = np.zeros((h + 8, w + 8))
    pad[4:-4, 4:-4] = plane
    tmp = np.zeros((h + 8, w))
    for d in range(9):
        tmp += k[d] * pad[:, d:d + w]
    out = np.zeros((h, w))
    for d in range(9):
        out += k[d] * tmp[d:d + h, :]
    return out


def test_iir_is_exactly_a_9_tap_fir():
    """Closed form: the double IIR's impulse response has support +-4 and unit gain."""
    import derive_iir

    taps, resp = derive_iir.fir_taps_f64()
    assert abs(resp[:16]).max() < 1e-15 and abs(resp[25:]).max() < 1e-15
    np.testing.assert_allclose(resp[16:25], resp[24:15:-1], atol=1e-15)  # symmetric
    assert abs(taps[0] + 2 * taps[1:].sum() - 1.0) < 1e-14
    # A Gaussian-like profile: monotone decreasing away from the centre.
    assert all(taps[i] > taps[i + 1] > 0 for i in range(4))


@pytest.mark.parametrize("shape", [(1, 1), (3, 2), (16, 16), (23, 41), (64, 9)])
def test_blur_f64_equals_brute_force_fir(oracle, shape):
    import derive_iir

    taps, _ = derive_iir.fir_taps_f64()
    rng = np.random.default_rng(7)
    p = rng.random(shape)
    np.testing.assert_allclose(oracle.plane_op("oracle_blur_f64", p), _fir_brute(p, taps), rtol=0, atol=1e-14)


@pytest.mark.parametrize("shape", [(32, 32), (17, 130), (200, 12), (256, 256)])
def test_blur_f32_iir_tracks_fir(oracle, shape):
    """The f32 IIR (cpu.rs forms) equals the exact FIR up to f32 drift along a line."""
    import derive_iir

    taps, _ = derive_iir.fir_taps_f64()
    rng = np.random.default_rng(0x5EED)
    p = rng.random(shape).astype(np.float32)
    ref = _fir_brute(p.astype(np.float64), taps)
    err = np.abs(oracle.plane_op("oracle_blur_f32", p) - ref).max()
    assert err < 1e-5, err  # f32 drift; a wrong constant or form gives ~1e-2
    # constant plane: interior is the constant (unit DC gain), edges fall off
    c = np.full(shape, 0.5, np.float32)
    b = oracle.plane_op("oracle_blur_f32", c)
    if min(shape) > 8:
        assert abs(b[4:-4, 4:-4] - 0.5).max() < 2e-6
        assert b[0, 0] < 0.5


def test_blur_delta_response_and_forms(oracle):
    """Delta at (8,12): the response is the outer product of the taps, and the
    row (MUL) and column (VERT) forms are each the 1-D FIR (blur_parity.rs cases)."""
    import derive_iir

    taps, _ = derive_iir.fir_taps_f64()
    p = np.zeros((24, 32), np.float32)
    p[8, 12] = 1.0
    k = np.concatenate([taps[:0:-1], taps])
    want = np.zeros((24, 32))
    want[4:13, 8:17] = np.outer(k, k)
    np.testing.assert_allclose(oracle.plane_op("oracle_blur_f32", p), want, atol=2e-7)
    rows = oracle.plane_op("oracle_blur_rows_f32", p)
    cols = oracle.plane_op("oracle_blur_cols_f32", p)
    np.testing.assert_allclose(rows[8, 8:17], k, atol=3e-7)
    np.testing.assert_allclose(cols[4:13, 12], k, atol=3e-7)
    # The two forms round differently: they are not the same function bit for bit.
    rng = np.random.default_rng(3)
    q = rng.random((40, 40)).astype(np.float32)
    assert not np.array_equal(oracle.plane_op("oracle_blur_rows_f32", q),
                              oracle.plane_op("oracle_blur_cols_f32", q.T.copy()).T)


# ---------------------------------------------------------------- XYB, downscale, scales


def test_xyb_f32_against_double_evaluation(oracle):
    """XYB in f32 vs the same closed form in double: within a few f32 ulps."""
    rng = np.random.default_rng(11)
    rgb = rng.random((5000, 3)).astype(np.float32)
    rgb[:8] = [[0, 0, 0], [1, 1, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0.5, 0.5, 0.5], [0, 1, 1], [1, 1, 0]]
    a = oracle.xyb_f32(rgb)
    b = oracle.xyb_f64(rgb.astype(np.float64))
    err = np.abs(a - b).max(axis=0)
    assert err[0] < 1.5e-6 and err[1] < 3e-7 and err[2] < 3e-7, err  # X carries a x14 gain
    # Known points: black maps to (0.42, 0.01, 0.55) exactly in double.
    np.testing.assert_allclose(b[0], [0.42, 0.01, 0.55], atol=1e-12)
    # Grey has no X opponent signal.
    assert abs(b[5, 0] - 0.42) < 1e-12


def test_cbrt_is_the_platform_cbrtf(oracle):
    """The f32 mirror uses the host cbrtf (as Rust f32::cbrt does); the GPU
    restatement of that method (ssim2_math.h cbrtf_ref) must match it. Here:
    the XYB of grey levels uses cbrtf(K + v), cross-checked with numpy cbrt of
    float32 where the two agree to 1 ulp."""
    g = np.linspace(0, 1, 4097, dtype=np.float32)
    rgb = np.stack([g, g, g], 1)
    y = oracle.xyb_f32(rgb)[:, 1]
    k = np.float32(0.0037930734)
    approx = (np.cbrt((k + g).astype(np.float64)) - 0.1559542025327239) + 0.01
    assert np.abs(y - approx).max() < 3e-7


def test_kernel_cbrt_equals_libm_on_every_float(tmp_path):
    """cbrtf_ref and cbrtf_nonneg (the GPU kernels' cube root, ssim2_math.h,
    with the exponent folded into a scale table) restate glibc's cbrtf, which
    cpu.rs reaches through f32::cbrt. Built here as host code and compared
    with the host cbrtf on every float in [0, 2] (the XYB argument
    range is [0.0038, 1.0038])."""
    import os
    import subprocess
    exe = tmp_path / "cbrt_exhaustive"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-std=c++17", "-pthread", "-D__HIP_PLATFORM_AMD__",
                    "-I/opt/rocm/include", f"-I{ROOT / 'turbo-metrics_amd' / 'csrc'}",
                    str(Path(__file__).parent / "cbrt_exhaustive.cpp"), "-o", str(exe), "-lm"], check=True)
    threads = str(max(1, min(8, os.cpu_count() or 1)))
    out = subprocess.run([str(exe), "0", "2", threads], check=True, capture_output=True, text=True).stdout
    checked, differ = map(int, out.split())
    assert checked == 0x40000000 + 1 and differ == 0


def test_downscale_brute_force(oracle):
    for h, w in [(1, 1), (2, 3), (5, 7), (16, 9), (33, 64)]:
        rng = np.random.default_rng(h * 100 + w)
        a = rng.random((h, w, 3)).astype(np.float32)
        oh, ow = (h + 1) // 2, (w + 1) // 2
        want = np.zeros((oh, ow, 3), np.float32)
        for oy in range(oh):
            for ox in range(ow):
                s = np.zeros(3, np.float32)
                for iy in range(2):
                    for ix in range(2):
                        s = s + a[min(2 * oy + iy, h - 1), min(2 * ox + ix, w - 1)]
                want[oy, ox] = s * np.float32(0.25)
        np.testing.assert_array_equal(oracle.downscale(a), want)


def test_scale_count_semantics(oracle):
    """cpu.rs:358-368: the <8 test runs before downscaling, so the first
    sub-8 scale is scored (Appendix semantics; differs from ssim2-gpu)."""
    assert oracle.scale_dims(64, 64) == [(64, 64), (32, 32), (16, 16), (8, 8), (4, 4)]
    assert oracle.scale_dims(16, 16) == [(16, 16), (8, 8), (4, 4)]
    assert oracle.scale_dims(8, 8) == [(8, 8), (4, 4)]
    assert oracle.scale_dims(7, 100) == []
    assert oracle.scale_dims(1920, 1080) == [(1920, 1080), (960, 540), (480, 270), (240, 135), (120, 68), (60, 34)]
    assert len(oracle.scale_dims(4000, 3000)) == 6


# ---------------------------------------------------------------- whole pipeline


def test_identical_inputs_score_exactly_100(oracle, synth):
    for kind, w, h in [(1, 64, 64), (2, 200, 150), (4, 123, 45)]:
        ref, _ = synth.pair(kind, 5, w, h)
        for mode in (MIRROR, F64, F32_FIR):
            assert oracle.score_srgb(ref, ref, mode) == 100.0


def test_too_small_scores_100(oracle):
    """No scale is scored below 8 px, so the weighted sum is 0 and the score 100."""
    a = np.zeros((7, 20, 3), np.uint8)
    b = np.full((7, 20, 3), 200, np.uint8)
    assert oracle.score_srgb(a, b) == 100.0


def test_score_monotone_in_distortion(oracle, synth):
    """parity_real_image.rs:93-119 pattern: stronger quantisation, lower score."""
    scores = []
    for step in (2, 8, 24, 64):
        ref, dis = synth.pair(synth.NOISE_GRAD_DCT, 77, 256, 256, step)
        scores.append(oracle.score_srgb(ref, dis))
    assert all(a > b for a, b in zip(scores, scores[1:])), scores


def test_weight_consumption_over_present_scales(oracle):
    """With fewer than 6 scales the weights are consumed sequentially over the
    present scales (cpu.rs:843-854): a 3-scale image's score equals the score
    of its sums laid out in scales 0..2 only, and moving the same sums to
    scales 3..5 changes it."""
    rng = np.random.default_rng(1)
    sums = np.zeros((6, 3, 6))
    sums[:3] = rng.random((3, 3, 6)) * 0.01
    dims = np.array([[16, 16], [8, 8], [4, 4], [0, 0], [0, 0], [0, 0]], np.int32)
    s3 = oracle.L.oracle_score_from_sums(sums.ctypes.data, dims.ctypes.data, 3)
    s6 = oracle.L.oracle_score_from_sums(
        np.concatenate([np.zeros((3, 3, 6)), sums[:3]]).ctypes.data,
        np.array([[16, 16], [8, 8], [4, 4], [16, 16], [8, 8], [4, 4]], np.int32).ctypes.data, 6)
    assert s3 != s6
    assert 0 < s3 < 100


def test_f32_mirror_vs_f64_gap_is_bounded(oracle, synth):
    """The mirror's own f32 error (the reference's) stays within the ~1-unit
    band measured on the workloads; the f32 FIR blur sits close to f64."""
    ref, dis = synth.pair(synth.NOISE_RECT_DCT, 2, 512, 384, 16)
    m, d, f = (oracle.score_srgb(ref, dis, mode) for mode in (MIRROR, F64, F32_FIR))
    assert abs(m - d) < 1.5
    assert abs(f - d) < 0.05


def test_jpeg_corpus_loose(oracle):
    """Published CPU ssimulacra2 0.5.1 scores for crates/dssim-cuda/test_data
    (ssim2-gpu/PORT_STATUS.md:28-33). The crate differs from cpu.rs in its sRGB
    transfer and cbrt, and the JPEGs are decoded here by PIL, so the check is
    loose; the corpus is read from tests/golden when present."""
    corpus = GOLDEN / "jpeg_corpus"
    if not (corpus / "source.png").exists():
        pytest.skip("JPEG corpus fixture not present")
    from PIL import Image

    src = np.asarray(Image.open(corpus / "source.png").convert("RGB"))
    published = {1: 1.2391, 5: -10.4452, 20: 57.0726, 45: 68.6823, 70: 79.5139, 90: 90.8900}
    for q, want in published.items():
        dis = np.asarray(Image.open(corpus / f"q{q}.jpg").convert("RGB"))
        got = oracle.score_srgb(src, dis)
        assert abs(got - want) < 0.25, (q, got, want)  # measured max 0.147 (q70)
