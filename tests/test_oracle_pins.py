"""Pins of the oracle's parts that the whole-pipeline checks leave open: the
binary16 round trip of the FP16-storage modes, the fast form's shifted
statistics (modes 7-9), the error maps and the score transform.

Each pin is an independent computation (numpy's IEEE binary16 conversion,
hand-worked pixels whose map values have a closed form, the score of a single
non-zero feature from cpu.rs:840-868's formula, the FP64 oracle for an exact
identity), and each is shown to fail when the matching plausible mistake is
planted in the oracle (oracle_plant).
"""
import json
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from conftest import GOLDEN

F64, F32_FIR, FAST, FAST_SHIFT, FAST_SHIFT_F16_ALL = 1, 2, 3, 7, 9
C2 = float(np.float32(0.0009))
TABLES = json.loads((GOLDEN / "reference_tables.json").read_text())
WEIGHT = np.array(TABLES["weight_f64_bits"], np.uint64).view(np.float64)


# ---------------------------------------------------------------- binary16

def _f16_mismatches(oracle, start, stop):
    bits = np.arange(start, stop, dtype=np.uint32)
    x = bits.view(np.float32)
    got = oracle.round_f16(x)
    want = x.astype(np.float16).astype(np.float32)
    bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
    neg = oracle.round_f16(-x)
    bad_neg = np.flatnonzero(neg.view(np.uint32) != (-want).view(np.uint32))
    return bad, bad_neg, x


def test_round_f16_equals_ieee_binary16_on_every_float(oracle):
    """Every float32 of magnitude below 65520 (the binary16 finite range after
    rounding), both signs, against numpy's IEEE round-to-nearest-even
    conversion: normals, subnormals, ties and zeros."""
    top = int(np.array(65520.0, np.float32).view(np.uint32))
    step = 1 << 22

    def chunk(start):
        bad, bad_neg, x = _f16_mismatches(oracle, start, min(start + step, top))
        return start, x[bad[:4]], x[bad_neg[:4]]

    with ThreadPoolExecutor(max(1, min(16, os.cpu_count() or 1))) as ex:
        for start, bad, bad_neg in ex.map(chunk, range(0, top, step)):
            assert bad.size == 0, (start, bad)
            assert bad_neg.size == 0, (start, bad_neg)


def test_round_f16_planted_truncation_is_caught(oracle):
    with oracle.planted(2):
        bad, _, _ = _f16_mismatches(oracle, 0x3F800000, 0x3F800000 + (1 << 16))
    assert bad.size > 0


# ------------------------------------------------------- shifted statistics

TINY = [(8, 8), (9, 13), (16, 16), (33, 17), (31, 33), (64, 64), (97, 150), (123, 45)]


def _tiny_pairs(synth):
    for i, (w, h) in enumerate(TINY):
        for kind, step in ((1, 16), (2, 8), (4, 40)):
            yield (w, h), synth.pair(kind, 300 + i, w, h, step)


def test_shifted_statistics_equal_f64(oracle, synth):
    """Mode 7 blurs x - c and rebuilds mean, variance and covariance with the
    zero-padding corrections (b1 = blur of ones, e = 1 - b1). The identities
    are exact, so in exact arithmetic mode 7 is the FP64 oracle (whose blur is
    the exact 9-tap FIR): only f32 rounding separates them. On tiny and odd
    images, where most pixels are within the blur's reach of a border and the
    corrections are large, the measured gap is 2.1e-4; without the corrections
    it is 16."""
    worst = 0.0
    for (w, h), (r, d) in _tiny_pairs(synth):
        worst = max(worst, abs(oracle.score_srgb(r, d, FAST_SHIFT) - oracle.score_srgb(r, d, F64)))
    assert worst <= 1e-3, worst
    # the same arithmetic without shifting (mode 3) loses more to cancellation
    # in f32 but is the same function
    (w, h), (r, d) = next(_tiny_pairs(synth))
    assert abs(oracle.score_srgb(r, d, FAST) - oracle.score_srgb(r, d, F64)) <= 1e-2


def test_shifted_statistics_planted_defect_is_caught(oracle, synth):
    (w, h), (r, d) = next(_tiny_pairs(synth))
    want = oracle.score_srgb(r, d, F64)
    with oracle.planted(1):
        assert abs(oracle.score_srgb(r, d, FAST_SHIFT) - want) > 0.1


def test_f16_storage_mode_bound(oracle, synth):
    """Mode 9 = mode 7 with every scale's XYB planes stored as binary16 of
    value - shift: within the FP16-storage bar (0.05) of mode 7 and of FP64."""
    worst7 = worst64 = 0.0
    for (w, h), (r, d) in _tiny_pairs(synth):
        s9 = oracle.score_srgb(r, d, FAST_SHIFT_F16_ALL)
        worst7 = max(worst7, abs(s9 - oracle.score_srgb(r, d, FAST_SHIFT)))
        worst64 = max(worst64, abs(s9 - oracle.score_srgb(r, d, F64)))
    assert worst7 <= 0.05 and worst64 <= 0.05, (worst7, worst64)
    assert worst7 > 0.0  # the storage rounding is applied


# ------------------------------------------------------------- error maps

def _ssim_d(mu1, mu2, v1, v2, cov):
    """cpu.rs:605-630 in exact arithmetic: 1 - num_m * num_s / den, clamped at 0 from below."""
    num_m = 1.0 - (mu1 - mu2) ** 2
    num_s = 2.0 * cov + C2
    den = v1 + v2 + C2
    return max(0.0, 1.0 - num_m * num_s / den)


# (mu1, mu2, var1, var2, cov): dyadic values, exact in f32
SSIM_CASES = [
    (0.5, 0.5, 0.0625, 0.015625, 0.03125),    # equal means: d from the variances only
    (0.5, 0.5, 0.03125, 0.03125, 0.03125),    # identical statistics: d = 0
    (0.5, 0.5, 0.03125, 0.03125, -0.03125),   # anti-correlated: d near 2 (no upper clamp)
    (0.75, 0.25, 0.0, 0.0, 0.0),              # flat, shifted mean: d = (mu1 - mu2)^2
    (0.625, 0.5, 0.015625, 0.0625, 0.0),      # both terms
    (0.5, 0.5, 0.0625, 0.0625, 0.0625 * 1.5), # cov > var: d clamped to 0
]


@pytest.mark.parametrize("fast", [False, True], ids=["mirror", "fast"])
def test_ssim_map_hand_worked(oracle, fast):
    mu1 = np.array([c[0] for c in SSIM_CASES], np.float32)
    mu2 = np.array([c[1] for c in SSIM_CASES], np.float32)
    s11 = np.array([c[0] ** 2 + c[2] for c in SSIM_CASES], np.float32)
    s22 = np.array([c[1] ** 2 + c[3] for c in SSIM_CASES], np.float32)
    s12 = np.array([c[0] * c[1] + c[4] for c in SSIM_CASES], np.float32)
    d = np.array([_ssim_d(*c) for c in SSIM_CASES])
    assert d[1] == 0.0 and d[2] > 1.9 and abs(d[3] - 0.25) < 1e-12 and d[5] == 0.0
    out = oracle.map_sums(mu1, mu2, mu1, mu2, s11, s22, s12, fast=fast)
    # src = mu1 and dis = mu2: no edge differences
    np.testing.assert_allclose(out[:2], [d.sum(), (d ** 4).sum()], rtol=2e-6)
    np.testing.assert_array_equal(out[2:], 0.0)
    with oracle.planted(4):
        bad = oracle.map_sums(mu1, mu2, mu1, mu2, s11, s22, s12, fast=fast)
    assert abs(bad[0] - d.sum()) > 0.1


@pytest.mark.parametrize("fast", [False, True], ids=["mirror", "fast"])
def test_edge_maps_hand_worked(oracle, fast):
    """cpu.rs:658-674: d1 = (1 + |dis - mu2|) / (1 + |src - mu1|) - 1,
    artifact = max(d1, 0), detail loss = max(-d1, 0)."""
    mu = np.array([0.5, 0.25, 0.75, 0.125], np.float32)
    delta = np.array([0.25, -0.125, 0.5, -0.0625], np.float32)
    flat = mu * mu  # zero variance and covariance: the SSIM term is 0
    # artifact only: src = mu1 and dis = mu2 + delta
    art = oracle.map_sums(mu, mu + delta, mu, mu, flat, flat, flat, fast=fast)
    a = np.abs(delta.astype(np.float64))
    np.testing.assert_allclose(art, [0, 0, a.sum(), (a ** 4).sum(), 0, 0], rtol=2e-6, atol=1e-12)
    # detail loss only: dis = mu2 and src = mu1 + delta
    det = oracle.map_sums(mu + delta, mu, mu, mu, flat, flat, flat, fast=fast)
    e = a / (1.0 + a)
    np.testing.assert_allclose(det, [0, 0, 0, 0, e.sum(), (e ** 4).sum()], rtol=2e-6, atol=1e-12)
    with oracle.planted(3):
        bad = oracle.map_sums(mu, mu + delta, mu, mu, flat, flat, flat, fast=fast)
    assert abs(bad[2] - a.sum()) > 0.1


# ------------------------------------------------------------------ score

def _score_closed_form(x):
    """cpu.rs:856-868 for a weighted feature sum x."""
    x *= 0.9562382616834844
    t = 6.248496625763138e-5 * x ** 3 - 0.020884521182843837 * x ** 2 + 2.326765642916932 * x
    return 100.0 - 10.0 * t ** 0.6276336467831387 if t > 0 else 100.0


def _single_feature(oracle, dims, c, s, n, k, f):
    """Score of sums holding one non-zero feature: channel c, scale s, norm n
    (0 = L1 mean, 1 = L4), map k (0 ssim, 1 artifact, 2 detail loss)."""
    ns = len(dims)
    sums = np.zeros((6, 3, 6))
    npx = dims[s][0] * dims[s][1]
    sums[s, c, 2 * k + n] = npx * (f if n == 0 else f ** 4)
    dd = np.zeros(12, np.int32)
    dd[:2 * ns] = np.array(dims).reshape(-1)
    return oracle.L.oracle_score_from_sums(sums.ctypes.data, dd.ctypes.data, ns)


def test_score_single_feature_closed_form(oracle):
    """Every one of the 108 features alone gives 100 - 10 t^0.6276... with t
    the cubic of 0.956... * WEIGHT[i] * feature, where i follows cpu.rs:843-854
    (channel, then present scale, then L1/L4, then ssim/artifact/detail loss)
    and the weights are the fixture extracted from cpu.rs."""
    dims = [(256, 256), (128, 128), (64, 64), (32, 32), (16, 16), (8, 8)]
    f = 0.05
    seen = 0
    for c in range(3):
        for s in range(6):
            for n in range(2):
                for k in range(3):
                    i = c * 36 + s * 6 + n * 3 + k
                    got = _single_feature(oracle, dims, c, s, n, k, f)
                    assert abs(got - _score_closed_form(WEIGHT[i] * f)) <= 1e-9, (c, s, n, k, got)
                    seen += WEIGHT[i] != 0
    assert seen > 50  # most features carry weight, so a permuted index shows


def test_score_weights_follow_present_scales(oracle):
    """With four scales the weights advance over the present scales only:
    channel c starts at index 24 c, not 36 c."""
    dims = [(64, 48), (32, 24), (16, 12), (8, 6)]
    for c, s, n, k in [(1, 0, 0, 0), (2, 3, 1, 2), (1, 2, 1, 1)]:
        i = c * 24 + s * 6 + n * 3 + k
        want = _score_closed_form(WEIGHT[i] * 0.05)
        assert abs(_single_feature(oracle, dims, c, s, n, k, 0.05) - want) <= 1e-9
    c, s, n, k = 1, 0, 0, 0
    assert WEIGHT[24] != WEIGHT[36]
    with oracle.planted(5):
        bad = _single_feature(oracle, dims, c, s, n, k, 0.05)
    assert abs(bad - _score_closed_form(WEIGHT[24] * 0.05)) > 1e-6
