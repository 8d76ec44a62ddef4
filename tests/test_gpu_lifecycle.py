"""Large-image lifecycle stress (crates/ssimulacra2-cuda/examples/drop_stress.rs:
31-128): fifteen sizes from 1024x768 to 5000x3500, five small-first passes;
per size a context is created, a reference is uploaded, converted with
srgb_to_linear and cached with set_reference_linear, three distorted images
are scored with compute_with_reference_linear, and everything is dropped.
Every pass must give the same scores, the cached-reference scores must equal
the direct ones, and the smallest size equals the oracle's f32 mirror."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [(1024, 768), (1600, 900), (2000, 1200), (2048, 1536), (1920, 1080), (2200, 1400), (2400, 1800),
         (2560, 1440), (2800, 1900), (3000, 2000), (3200, 2100), (3500, 2333), (4000, 3000), (4500, 3000),
         (5000, 3500)]


def _say(msg):
    print(msg, flush=True)


def make_rgb(w, h, seed):
    """drop_stress.rs make_rgb: deterministic gradients with u8 wrap-around."""
    x = np.arange(w, dtype=np.int64)[None, :]
    y = np.arange(h, dtype=np.int64)[:, None]
    out = np.empty((h, w, 3), np.uint8)
    out[..., 0] = (((x + seed) & 255) * 3) & 255
    out[..., 1] = (((y + seed) & 255) * 5) & 255
    out[..., 2] = ((((x ^ y) & 255) + seed) * 7) & 255
    return out


@pytest.fixture(scope="module")
def s2():
    import ssimulacra2_hip
    return ssimulacra2_hip


def test_drop_stress_sizes(s2, oracle):
    scores = {}
    for p in range(5):
        for w, h in SIZES:
            _say(f"[gpu] lifecycle pass {p} {w}x{h}")
            ref_host = make_rgb(w, h, 0)
            dis_host = [make_rgb(w, h, k + 1) for k in range(3)]
            stream = s2.Stream()
            ref_dev, dis_dev = s2.DeviceImage.from_host(ref_host), s2.DeviceImage(w, h)
            ref_lin, dis_lin = s2.DeviceImage(w, h, np.float32), s2.DeviceImage(w, h, np.float32)
            m = s2.Ssimulacra2.new(ref_lin, dis_lin, stream)
            k = m.kernel()
            k.srgb_to_linear(ref_dev, ref_lin, stream)
            m.set_reference_linear(ref_lin, stream)
            got = []
            for d in dis_host:
                dis_dev.upload(d)
                k.srgb_to_linear(dis_dev, dis_lin, stream)
                got.append(m.compute_with_reference_linear(ref_lin, dis_lin, stream))
            if p == 0:
                direct = []
                for d in dis_host:
                    dis_dev.upload(d)
                    k.srgb_to_linear(dis_dev, dis_lin, stream)
                    direct.append(m.compute_sync(stream))
                assert got == direct, ((w, h), got, direct)
                scores[(w, h)] = got
            else:
                assert got == scores[(w, h)], ((w, h), p, got, scores[(w, h)])
            m.close()
            for im in (ref_dev, dis_dev, ref_lin, dis_lin):
                im.free()
            stream.close()
    w, h = SIZES[0]
    want = oracle.score_srgb(make_rgb(w, h, 0), make_rgb(w, h, 1))
    _say(f"[gpu] {w}x{h}: gpu {scores[(w, h)][0]:.9f} oracle {want:.9f}")
    assert abs(scores[(w, h)][0] - want) <= 1e-6
    assert all(np.isfinite(v) for s in scores.values() for v in s)
