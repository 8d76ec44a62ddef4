"""GPU test of ssim2_compute_srgb_noalloc_async, the entry point that allocates
nothing: called through the C ABI, its result record must equal what the
context path (ssim2_compute_srgb_async + get_score / get_sums) gives for the
same pair, bit for bit, on consecutive calls whose second pair is smaller (the
first call leaves its planes where the second one's padding must read zero).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RESULT_DOUBLES = 121


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def s2():
    import ssimulacra2_hip
    return ssimulacra2_hip


def _context_record(s2, ref, dis):
    """Score, sums and dims of the context path for one device pair."""
    with s2.Ssimulacra2(ref.width, ref.height) as m:
        m.compute_srgb(ref, dis)
        score = m.get_score()
        sums, dims, n = m.sums()
    return score, sums.reshape(-1), dims.reshape(-1), n


def _noalloc(s2, ref, dis, out, stream=None):
    st = s2.lib.ssim2_compute_srgb_noalloc_async(ref.as_struct(), dis.as_struct(), out.ptr,
                                                  stream.handle if stream else None)
    assert st == 0, s2.lib.ssim2_last_error()
    if stream:
        stream.sync()
    else:
        s2.device_synchronize()
    return out.download(np.float64)


def test_no_alloc_equals_context_path(s2, synth):
    # (kind, seed, w, h, step): a large pair, then smaller ones of other pitches
    cases = [(2, 71, 1920, 1080, 16), (4, 72, 777, 555, 8), (1, 73, 64, 64, 16), (2, 74, 8, 8, 16)]
    out = s2.DeviceBuffer(RESULT_DOUBLES * 8)
    stream = s2.Stream()
    for k, (kind, seed, w, h, step) in enumerate(cases):
        r, d = synth.pair(kind, seed, w, h, step)
        ref, dis = s2.DeviceImage.from_host(r), s2.DeviceImage.from_host(d)
        _say(f"[gpu] no-alloc call {k}: {w}x{h}")
        rec = _noalloc(s2, ref, dis, out, stream if k % 2 else None)
        score, sums, dims, n = _context_record(s2, ref, dis)
        _say(f"[gpu] {w}x{h}: no-alloc {rec[0]:.12f} context {score:.12f}")
        assert rec[0] == score
        np.testing.assert_array_equal(rec[1:109], sums)
        np.testing.assert_array_equal(rec[109:121].astype(np.int32), dims)
    # the same large pair again after the small ones
    r, d = synth.pair(*cases[0][:2], *cases[0][2:4], cases[0][4])
    ref, dis = s2.DeviceImage.from_host(r), s2.DeviceImage.from_host(d)
    rec = _noalloc(s2, ref, dis, out)
    assert rec[0] == _context_record(s2, ref, dis)[0]
    stream.close()


def test_no_alloc_refuses_beyond_its_scratch(s2):
    big = s2.DeviceImage(2048, 1152)
    out = s2.DeviceBuffer(RESULT_DOUBLES * 8)
    st = s2.lib.ssim2_compute_srgb_noalloc_async(big.as_struct(), big.as_struct(), out.ptr, None)
    assert s2.STATUS[st] == "SSIM2_INVALID_SIZE"
    assert b"scratch" in s2.lib.ssim2_last_error()
    small = s2.DeviceImage(7, 100)
    assert s2.STATUS[s2.lib.ssim2_compute_srgb_noalloc_async(small.as_struct(), small.as_struct(), out.ptr,
                                                             None)] == "SSIM2_INVALID_SIZE"
    other = s2.DeviceImage(64, 64)
    assert s2.STATUS[s2.lib.ssim2_compute_srgb_noalloc_async(big.as_struct(), other.as_struct(), out.ptr,
                                                             None)] == "SSIM2_DIM_MISMATCH"
    assert s2.STATUS[s2.lib.ssim2_compute_srgb_noalloc_async(other.as_struct(), other.as_struct(), None,
                                                             None)] == "SSIM2_INVALID_ARGUMENT"
