"""CPU tests of ssim2_stats_compute, the restatement of
quick_stats::full::Stats::compute (crates/quick-stats/src/lib.rs:21-98).
Pinned by the reference's own sample (lib.rs:101-107, values 0, 1, 3, 4, worked
by hand here), by numpy's independent percentile (linear interpolation) and
variance, and by the special cases of the definition."""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def video():
    from ssimulacra2_hip import video as v
    return v


def test_reference_sample(video):
    st = video.Stats.compute([0.0, 1.0, 3.0, 4.0])
    # mean 2; squared deviations 4, 1, 1, 4 -> 10; /4 and /3
    assert (st.min, st.max, st.mean) == (0.0, 4.0, 2.0)
    assert st.var == 2.5 and st.sample_var == 10.0 / 3.0
    assert st.stddev == math.sqrt(2.5) and st.sample_stddev == math.sqrt(10.0 / 3.0)
    # rank = pct/100 * 3: p50 -> 1.5 -> 1 + (3-1)*0.5; p1 -> 0.03; p95 -> 2.85 -> 3 + 0.85
    assert st.p50 == 2.0
    assert abs(st.p1 - 0.03) < 1e-15 and abs(st.p95 - 3.85) < 1e-15


def test_against_numpy(video):
    rng = np.random.default_rng(11)
    for n in (2, 3, 17, 240, 1001):
        v = rng.normal(80, 7, n)
        st = video.Stats.compute(v)
        assert st.min == v.min() and st.max == v.max()
        assert abs(st.mean - v.mean()) < 1e-12
        assert abs(st.var - v.var()) < 1e-9 and abs(st.sample_var - v.var(ddof=1)) < 1e-9
        for p in (1, 5, 50, 95, 99):
            assert abs(getattr(st, f"p{p}") - np.percentile(v, p)) < 1e-9, (n, p)


def test_single_value_and_order(video):
    st = video.Stats.compute([42.5])
    assert st.min == st.max == st.mean == st.p1 == st.p99 == 42.5
    assert st.var == 0.0 and st.sample_var == 0.0
    # f64::total_cmp order: -0.0 sorts before +0.0
    st = video.Stats.compute([0.0, -0.0])
    assert math.copysign(1.0, st.min) == -1.0 and math.copysign(1.0, st.max) == 1.0


def test_empty_is_an_error(video):
    import ssimulacra2_hip as s2
    with pytest.raises(s2.Ssim2Error):
        video.Stats.compute([])


def test_output_formats(video):
    agg = video.MetricAggregate([90.5, 80.25], video.Stats.compute([90.5, 80.25]))
    r = video.MetricsResults(2, agg)
    import json
    d = json.loads(r.to_json())
    assert d["frame_count"] == 2 and d["ssimulacra2"]["scores"] == [90.5, 80.25]
    assert r.to_csv().splitlines() == ["ssimulacra2", "90.5", "80.25"]
    lines = r.to_json_lines().splitlines()
    assert json.loads(lines[0]) == {"ssimulacra2": 90.5} and json.loads(lines[-1])["frame_count"] == 2
    assert r.to_default().startswith("SSIMULACRA2: Stats {")
