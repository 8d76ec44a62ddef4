"""GPU tests of the video driver (ssim2_video_compute_all), the counterpart of
TurboMetrics::compute_all (crates/turbo-metrics/src/lib.rs:367-442): frame
sampling (skip, skip_ref, skip_dis, every, frames) and that every kept pair
scores exactly as it does through the direct entry points."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W, H = 96, 64


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def s2():
    import ssimulacra2_hip
    return ssimulacra2_hip


@pytest.fixture(scope="module")
def video():
    from ssimulacra2_hip import video as v
    return v


@pytest.fixture(scope="module")
def yuv_frames(s2, synth):
    ref, dis = [], []
    for t in range(9):
        r, d = synth.yuv420_pair(31, W, H, t)
        ref.append(s2.DeviceYuv.from_planes(*r))
        dis.append(s2.DeviceYuv.from_planes(*d))
    return ref, dis


def kept(n_ref, n_dis, o):
    """(ref index, dis index) of the scored pairs, by the rules of lib.rs:388-405."""
    i, j = o.skip + o.skip_ref, o.skip + o.skip_dis
    out, decoded = [], 0
    while i < n_ref and j < n_dis:
        k = decoded
        if o.every > 1 and k != 0 and k % o.every != 0:
            decoded += 1
        elif o.frames > 0 and k >= o.frames:
            break
        else:
            decoded += 1
            out.append((i, j))
        i += 1
        j += 1
    return out


OPTS = [dict(), dict(every=2, skip=1), dict(skip_ref=2, skip_dis=1, frames=3), dict(every=3, frames=7)]


@pytest.mark.parametrize("opt", OPTS)
def test_exact_mode_equals_direct(s2, video, yuv_frames, opt):
    ref, dis = yuv_frames
    o = video.Options(**opt)
    _say(f"[gpu] video exact {opt}")
    res = video.compute_all(W, H, ref, dis, o, mode="exact")
    pairs = kept(len(ref), len(dis), o)
    with s2.Ssimulacra2(W, H) as m:
        want = [m.compute_yuv420_sync(ref[i], dis[j]) for i, j in pairs]
    assert res.frame_count == len(pairs)
    assert res.ssimulacra2.scores == want
    assert res.ssimulacra2.stats == video.Stats.compute(want)


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_fast_mode_equals_batch(s2, video, yuv_frames, storage):
    ref, dis = yuv_frames
    o = video.Options(every=2)
    _say(f"[gpu] video fast {storage}")
    res = video.compute_all(W, H, ref, dis, o, mode="fast", storage=storage, batch=2)  # flushes mid-stream
    pairs = kept(len(ref), len(dis), o)
    with s2.Ssim2Batch(max_pairs=8, max_pixels=8 * W * H, storage=storage) as b:
        want = list(b.compute_yuv([(ref[i], dis[j]) for i, j in pairs]))
    assert res.ssimulacra2.scores == want


def test_srgb_frames(s2, video, synth):
    ref, dis = [], []
    for t in range(4):
        r, d = synth.pair(synth.NOISE_GRAD_DCT, 40 + t, W, H, 8 + 8 * t)
        ref.append(s2.DeviceImage.from_host(r))
        dis.append(s2.DeviceImage.from_host(d))
    _say("[gpu] video srgb")
    exact = video.compute_all(W, H, ref, dis, mode="exact")
    fast = video.compute_all(W, H, ref, dis, mode="fast", storage="f32", batch=3)
    with s2.Ssimulacra2(W, H) as m:
        want = [m.compute_srgb_sync(r, d) for r, d in zip(ref, dis)]
    assert exact.ssimulacra2.scores == want
    assert np.max(np.abs(np.array(fast.ssimulacra2.scores) - want)) < 0.05


def test_size_mismatch_is_an_error(s2, video, yuv_frames):
    ref, dis = yuv_frames
    _say("[gpu] video size mismatch")
    with pytest.raises(s2.Ssim2Error) as ei:
        video.compute_all(W, H + 2, ref, dis)
    assert ei.value.status == "SSIM2_DIM_MISMATCH"
