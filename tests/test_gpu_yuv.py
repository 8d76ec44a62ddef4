"""GPU tests of the video input: YUV 4:2:0 10-bit BT.709 limited range
(crates/turbo-metrics/src/lib.rs:271-363 converts each decoded frame with
crates/cuda-colorspace-kernel/src/biplanar.rs:8-70, then scores it).

Bars: conversion within 2e-6 of the oracle's double conversion (f32
coefficients and the device powf against double); the fused exact path is
bit-identical to converting first and scoring the linear pair, and equals the
oracle's f32 mirror on the converted planes to 1e-6; P016 and planar 10-bit
give identical scores; the FP16-storage fast form is within 0.05 of the FP64
oracle, which converts and scores in double."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MIRROR = 0
BAR_F64 = 0.05


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def s2():
    import ssimulacra2_hip
    return ssimulacra2_hip


SIZES = [(97, 61), (256, 144), (64, 8), (333, 200)]


@pytest.mark.parametrize("w,h", SIZES)
def test_conversion_matches_oracle(s2, oracle, synth, w, h):
    ref, dis = synth.yuv420_pair(21, w, h, 3)
    _say(f"[gpu] yuv420_to_linear {w}x{h}")
    want = oracle.yuv_to_linear(*dis)
    out10 = s2.DeviceImage(w, h, np.float32)
    s2.yuv420_to_linear(s2.DeviceYuv.from_planes(*dis), out10)
    out16 = s2.DeviceImage(w, h, np.float32)
    s2.yuv420_to_linear(s2.DeviceYuv.from_planes(*dis, fmt="p016"), out16)
    a, b = out10.download(), out16.download()
    assert np.array_equal(a, b)
    assert np.max(np.abs(a.astype(np.float64) - want)) <= 2e-6


@pytest.mark.parametrize("w,h", SIZES)
def test_exact_path_fused_equals_convert_then_score(s2, oracle, synth, w, h):
    ref, dis = synth.yuv420_pair(22, w, h, 1)
    _say(f"[gpu] exact yuv {w}x{h}")
    R, D = s2.DeviceYuv.from_planes(*ref), s2.DeviceYuv.from_planes(*dis)
    lr, ld = s2.DeviceImage(w, h, np.float32), s2.DeviceImage(w, h, np.float32)
    s2.yuv420_to_linear(R, lr)
    s2.yuv420_to_linear(D, ld)
    with s2.Ssimulacra2(w, h) as m:
        fused = m.compute_yuv420_sync(R, D)
        linear = m.compute_sync(ref_linear=lr, dis_linear=ld)
        p016 = m.compute_yuv420_sync(s2.DeviceYuv.from_planes(*ref, fmt="p016"),
                                     s2.DeviceYuv.from_planes(*dis, fmt="p016"))
    assert fused == linear == p016
    a, b = lr.download(), ld.download()
    mirror = oracle.score_linear(a, b, MIRROR)
    f64 = oracle.score_yuv_f64(ref, dis)
    _say(f"[gpu] {w}x{h}: gpu {fused:.6f} mirror {mirror:.6f} f64 {f64:.6f}")
    assert abs(fused - mirror) <= 1e-6


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_batched_fast_form_vs_f64(s2, oracle, synth, storage):
    frames = [(synth.yuv420_pair(23, w, h, t), (w, h)) for t, (w, h) in enumerate(SIZES + [(640, 360)])]
    dev = [(s2.DeviceYuv.from_planes(*r), s2.DeviceYuv.from_planes(*d)) for (r, d), _ in frames]
    _say(f"[gpu] batch yuv {storage}, {len(dev)} frames")
    with s2.Ssim2Batch(max_pairs=16, max_pixels=2_000_000, storage=storage) as b:
        together = b.compute_yuv(dev)
        alone = [b.compute_yuv([p])[0] for p in dev]
        p016 = b.compute_yuv([(s2.DeviceYuv.from_planes(*r, fmt="p016"), s2.DeviceYuv.from_planes(*d, fmt="p016"))
                              for (r, d), _ in frames])
    np.testing.assert_array_equal(together, np.array(alone))
    np.testing.assert_array_equal(together, p016)
    for ((r, d), (w, h)), sc in zip(frames, together):
        f64 = oracle.score_yuv_f64(r, d)
        _say(f"[gpu] {storage} {w}x{h}: {sc:.6f} f64 {f64:.6f}")
        assert abs(sc - f64) <= BAR_F64, (w, h, sc, f64)


def test_yuv_errors(s2, synth):
    ref, _ = synth.yuv420_pair(24, 64, 64, 0)
    a = s2.DeviceYuv.from_planes(*ref)
    b = s2.DeviceYuv.from_planes(*ref, fmt="p016")
    small = s2.DeviceYuv(64, 32)
    with s2.Ssimulacra2(64, 64) as m:
        with pytest.raises(s2.Ssim2Error) as ei:
            m.compute_yuv420_sync(a, b)
        assert ei.value.status == "SSIM2_INVALID_ARGUMENT"
        with pytest.raises(s2.Ssim2Error) as ei:
            m.compute_yuv420_sync(a, small)
        assert ei.value.status == "SSIM2_DIM_MISMATCH"
        assert m.compute_yuv420_sync(a, a) == 100.0
