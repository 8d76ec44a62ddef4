"""The synthetic workload generators are deterministic and shaped as specified."""
import json

import numpy as np

from conftest import GOLDEN


def test_digests_are_stable(synth):
    import importlib.util
    spec = importlib.util.spec_from_file_location("mk", GOLDEN / "make_synth_hashes.py")
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    assert mk.digests() == json.loads((GOLDEN / "synth_sha256.json").read_text())


def test_pair_properties(synth):
    ref, dis = synth.pair(synth.NOISE_RECT_DCT, 2, 128, 64, 16)
    assert ref.shape == dis.shape == (64, 128, 3) and ref.dtype == np.uint8
    assert 0 < np.abs(ref.astype(int) - dis).mean() < 20
    r2, d2 = synth.pair(synth.NOISE_RECT_DCT, 2, 128, 64, 16)
    assert np.array_equal(ref, r2) and np.array_equal(dis, d2)
    r3, _ = synth.pair(synth.NOISE_RECT_DCT, 3, 128, 64, 16)
    assert not np.array_equal(ref, r3)
    # coarser quantisation moves further from the reference
    _, dq = synth.pair(synth.NOISE_GRAD_DCT, 9, 128, 128, 48)
    _, df = synth.pair(synth.NOISE_GRAD_DCT, 9, 128, 128, 4)
    rr, _ = synth.pair(synth.NOISE_GRAD_DCT, 9, 128, 128, 4)
    assert np.abs(rr.astype(int) - dq).mean() > np.abs(rr.astype(int) - df).mean()


def test_log_uniform_range(synth):
    v = [synth.log_uniform(11, i, 1, 64, 2048) for i in range(2000)]
    assert min(v) >= 64 and max(v) <= 2048
    # log-uniform: about half below the geometric mean sqrt(64*2048) = 362
    frac = np.mean(np.array(v) < 362)
    assert 0.42 < frac < 0.58


def test_yuv420_ranges(synth):
    (y, u, v), (yd, ud, vd) = synth.yuv420_pair(3, 64, 32, 0)
    assert y.shape == (32, 64) and u.shape == (16, 32)
    assert y.min() >= 64 and y.max() <= 940 and u.min() >= 64 and u.max() <= 960
    assert set(np.unique((yd.astype(int) - 64) % 16)) == {0}
