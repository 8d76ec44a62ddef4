"""GPU tests of the serving runtime (ssim2_server_*): continuous batching,
admission by HBM budget, and that every result is the same alone and in the
mix (bit-identical to scoring the pair by itself)."""
import random
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def s2():
    import ssimulacra2_hip
    return ssimulacra2_hip


@pytest.fixture(scope="module")
def pool(s2, synth):
    sizes = [(256, 256), (320, 180), (64, 64), (97, 203), (640, 360), (16, 16), (1024, 96), (128, 512)]
    out = []
    for i, (w, h) in enumerate(sizes):
        r, d = synth.pair(synth.NOISE_GRAD_DCT, 500 + i, w, h, 4 + 4 * (i % 6))
        out.append((s2.DeviceImage.from_host(r), s2.DeviceImage.from_host(d)))
    return out


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_server_results_equal_alone(s2, pool, storage):
    with s2.Ssim2Batch(max_pairs=8, max_pixels=2_000_000, storage=storage) as b:
        alone = [b.compute([p])[0] for p in pool]
    rng = random.Random(7)
    order = [rng.randrange(len(pool)) for _ in range(60)]
    _say(f"[gpu] server {storage}: 60 requests over {len(pool)} pairs")
    with s2.Ssim2Server(storage=storage, slots=3, max_pairs_per_batch=16, max_pixels_per_batch=4_000_000,
                        hbm_budget=1 << 30) as sv:
        for rid, k in enumerate(order):
            sv.submit(*pool[k], rid)
            if rid % 7 == 0:
                time.sleep(0.0005)  # let batches start while more requests arrive
        sv.drain(60.0)
        res = sv.poll()
        st = sv.stats()
    assert len(res) == len(order)
    assert st["completed"] == len(order) and st["refused"] == 0 and st["batches"] >= 2
    for rid, score, t0, t1, status in res:
        assert status == "SSIM2_OK"
        assert t1 >= t0
        assert score == alone[order[rid]], (rid, score, alone[order[rid]])


@pytest.mark.parametrize("storage", ["f16", "f32"])
def test_server_resident_references(s2, pool, storage):
    """Requests naming a resident reference batch with direct ones and score
    bit-identically to the direct pair."""
    with s2.Ssim2Batch(max_pairs=8, max_pixels=2_000_000, storage=storage) as b:
        alone = [b.compute([p])[0] for p in pool]
    rng = random.Random(11)
    order = [(rng.randrange(len(pool)), rng.random() < 0.6) for _ in range(80)]
    _say(f"[gpu] server {storage}: 80 requests, resident references for 60%")
    with s2.Ssim2Server(storage=storage, slots=3, max_pairs_per_batch=16, max_pixels_per_batch=4_000_000,
                        hbm_budget=1 << 30) as sv:
        ids = [sv.add_reference(r) for r, _ in pool]
        for rid, (k, by_ref) in enumerate(order):
            if by_ref:
                sv.submit_ref(ids[k], pool[k][1], rid)
            else:
                sv.submit(*pool[k], rid)
            if rid % 9 == 0:
                time.sleep(0.0005)
        sv.drain(60.0)
        res = sv.poll()
        assert len(res) == len(order)
        for rid, score, t0, t1, status in res:
            assert status == "SSIM2_OK"
            assert score == alone[order[rid][0]], (rid, score, alone[order[rid][0]])
        sv.remove_reference(ids[0])
        with pytest.raises(s2.Ssim2Error) as ei:
            sv.submit_ref(ids[0], pool[0][1], 1000)
        assert ei.value.status == "SSIM2_NO_REFERENCE"
        with pytest.raises(s2.Ssim2Error) as ei:
            sv.submit_ref(ids[1], pool[0][1], 1001)
        assert ei.value.status == "SSIM2_DIM_MISMATCH"
        sv.submit_ref(ids[1], pool[1][1], 1002)  # the others stay resident
        sv.drain(60.0)
        (rid, score, _, _, status), = sv.poll()
        assert rid == 1002 and status == "SSIM2_OK" and score == alone[1]


def test_server_admission(s2, pool):
    big = s2.DeviceImage(700, 700)
    with s2.Ssim2Server(storage="f16", slots=1, max_pairs_per_batch=4, max_pixels_per_batch=300_000,
                        hbm_budget=1 << 30) as sv:
        with pytest.raises(s2.Ssim2Error) as ei:
            sv.submit(big, big, 1)
        assert ei.value.status == "SSIM2_ADMISSION_REFUSED"
        assert sv.stats()["refused"] == 1


def test_server_tight_budget_serialises(s2, pool):
    """A budget that holds one 640x360 request at a time still completes all."""
    with s2.Ssim2Batch(max_pairs=4, max_pixels=1_000_000, storage="f16") as b:
        want = b.compute([pool[4]])[0]
    _say("[gpu] server tight budget")
    with s2.Ssim2Server(storage="f16", slots=2, max_pairs_per_batch=4, max_pixels_per_batch=1_000_000,
                        hbm_budget=5_000_000) as sv:
        for rid in range(6):
            sv.submit(*pool[4], rid)
        sv.drain(60.0)
        res = sv.poll()
        st = sv.stats()
    assert [r[1] for r in res] == [want] * 6
    assert st["batches"] >= 3


def test_server_remove_reference_while_queued(s2, pool):
    """A resident reference removed while requests naming it wait: those
    requests fail alone with SSIM2_NO_REFERENCE (or score normally if they ran
    before the removal), and every other request of their batches scores as
    it does alone."""
    with s2.Ssim2Batch(max_pairs=8, max_pixels=2_000_000, storage="f16") as b:
        alone = [b.compute([p])[0] for p in pool]
    _say("[gpu] server: remove a reference with 150 of its requests queued")
    with s2.Ssim2Server(storage="f16", slots=1, max_pairs_per_batch=2, max_pixels_per_batch=4_000_000,
                        hbm_budget=1 << 30) as sv:
        ref_id = sv.add_reference(pool[4][0])
        kinds = {}
        for rid in range(300):
            if rid % 2 == 0:
                sv.submit_ref(ref_id, pool[4][1], rid)
                kinds[rid] = None
            else:
                k = rid % len(pool)
                sv.submit(*pool[k], rid)
                kinds[rid] = k
        sv.remove_reference(ref_id)
        for rid in range(300, 320):
            k = rid % len(pool)
            sv.submit(*pool[k], rid)
            kinds[rid] = k
        sv.drain(60.0)
        res = sv.poll()
    assert len(res) == len(kinds)
    stale = 0
    for rid, score, _, _, status in res:
        k = kinds[rid]
        if k is None:
            assert status in ("SSIM2_OK", "SSIM2_NO_REFERENCE"), (rid, status)
            if status == "SSIM2_OK":
                assert score == alone[4], (rid, score, alone[4])
            else:
                stale += 1
        else:
            assert status == "SSIM2_OK", (rid, status)
            assert score == alone[k], (rid, score, alone[k])
    _say(f"[gpu] {stale} of 150 reference requests were still queued at removal")
    assert stale >= 1
