"""The video/metric driver for ssimulacra2 and its statistics: the Python face
of ssim2_video_compute_all and ssim2_stats_compute.

Mirrors crates/turbo-metrics/src/lib.rs (Options, MetricAggregate,
MetricsResults, TurboMetrics::compute_all), crates/quick-stats/src/lib.rs
(Stats) and the output formats of crates/turbo-metrics-cli/src/output.rs
(default, json, json-lines, csv) for the ssimulacra2 metric.
"""
from __future__ import annotations

import ctypes
import json
from dataclasses import asdict, dataclass, field
from typing import Iterable, Iterator, List, Optional

import numpy as np

from . import (FRAME_SOURCE, STORAGE, DeviceImage, DeviceYuv, Ssim2Error, _check, _Frame, _Stats, _stream_ptr,
               _VideoOptions, lib)

_STATS_FIELDS = ("min", "max", "mean", "var", "sample_var", "stddev", "sample_stddev", "p1", "p5", "p50", "p95",
                 "p99")


@dataclass
class Stats:
    """quick_stats::full::Stats (crates/quick-stats/src/lib.rs:1-19)."""
    min: float
    max: float
    mean: float
    var: float
    sample_var: float
    stddev: float
    sample_stddev: float
    p1: float
    p5: float
    p50: float
    p95: float
    p99: float

    @classmethod
    def _from_c(cls, st: _Stats) -> "Stats":
        return cls(**{k: getattr(st, k) for k in _STATS_FIELDS})

    @classmethod
    def compute(cls, values) -> "Stats":
        """Stats::compute (quick-stats lib.rs:21-52), in the native library."""
        v = np.ascontiguousarray(values, dtype=np.float64)
        st = _Stats()
        _check(lib.ssim2_stats_compute(v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), v.size,
                                       ctypes.byref(st)))
        return cls._from_c(st)


@dataclass
class Options:
    """turbo_metrics::Options (crates/turbo-metrics/src/lib.rs:44-59)."""
    every: int = 0
    skip: int = 0
    skip_ref: int = 0
    skip_dis: int = 0
    frames: int = 0


@dataclass
class MetricAggregate:
    """lib.rs:61-77: per-frame scores and their statistics."""
    scores: List[float]
    stats: Stats


@dataclass
class MetricsResults:
    """lib.rs:79-91, for the ssimulacra2 metric."""
    frame_count: int
    ssimulacra2: Optional[MetricAggregate] = field(default=None)

    def to_json(self) -> str:
        """Output::Json (output.rs): the whole object, pretty-printed."""
        d = {"frame_count": self.frame_count}
        if self.ssimulacra2 is not None:
            d["ssimulacra2"] = {"scores": self.ssimulacra2.scores, "stats": asdict(self.ssimulacra2.stats)}
        return json.dumps(d, indent=2)

    def to_json_lines(self) -> str:
        """Output::JsonLines: one object per frame, then the statistics."""
        lines = [json.dumps({"ssimulacra2": s}) for s in (self.ssimulacra2.scores if self.ssimulacra2 else [])]
        tail = {"frame_count": self.frame_count}
        if self.ssimulacra2 is not None:
            tail["ssimulacra2"] = asdict(self.ssimulacra2.stats)
        lines.append(json.dumps(tail))
        return "\n".join(lines)

    def to_csv(self) -> str:
        """Output::CSV: a header, then one score per frame."""
        rows = ["ssimulacra2"] + [repr(float(s)) for s in (self.ssimulacra2.scores if self.ssimulacra2 else [])]
        return "\n".join(rows) + "\n"

    def to_default(self) -> str:
        """Output::Default: the statistics for a human reader."""
        if self.ssimulacra2 is None:
            return ""
        body = "\n".join(f"    {k}: {v}," for k, v in asdict(self.ssimulacra2.stats).items())
        return f"SSIMULACRA2: Stats {{\n{body}\n}}"


def _frame_of(obj, out: _Frame):
    if isinstance(obj, DeviceYuv):
        out.kind = 1
        out.yuv = obj._img
    elif isinstance(obj, DeviceImage):
        if obj.dtype != np.uint8:
            raise Ssim2Error(8, "frames are sRGB u8 images or YUV 4:2:0 frames")
        out.kind = 0
        out.srgb = obj._img
    else:
        raise TypeError(f"unsupported frame type {type(obj)!r}")


def _source(frames: Iterable):
    it: Iterator = iter(frames)
    keep = []  # the current frame stays alive until the next pull

    def pull(_user, frame_ptr):
        try:
            obj = next(it)
        except StopIteration:
            return 0
        try:
            keep[:] = [obj]
            _frame_of(obj, frame_ptr.contents)
            return 1
        except Exception:
            return -1

    return FRAME_SOURCE(pull)


def compute_all(width: int, height: int, frames_ref: Iterable, frames_dis: Iterable,
                options: Options = Options(), mode: str = "exact", storage: str = "f16", batch: int = 8,
                stream=None, capacity: int = 1 << 20) -> MetricsResults:
    """TurboMetrics::compute_all for ssimulacra2 (lib.rs:367-442). Frames are
    DeviceImage (sRGB u8) or DeviceYuv objects; `mode` is "exact" (the exact
    path, pair by pair) or "fast" (the batched fast form with `storage`)."""
    o = _VideoOptions()
    o.struct_size = ctypes.sizeof(_VideoOptions)
    o.every, o.skip, o.skip_ref, o.skip_dis, o.frames = (options.every, options.skip, options.skip_ref,
                                                          options.skip_dis, options.frames)
    o.mode = {"exact": 0, "fast": 1}[mode]
    o.storage = STORAGE[storage]
    o.batch = batch
    src_r, src_d = _source(frames_ref), _source(frames_dis)
    scores = (ctypes.c_double * capacity)()
    n = ctypes.c_size_t()
    st = _Stats()
    _check(lib.ssim2_video_compute_all(width, height, ctypes.byref(o), src_r, None, src_d, None,
                                       _stream_ptr(stream), scores, capacity, ctypes.byref(n), ctypes.byref(st)))
    vals = [float(x) for x in scores[:min(n.value, capacity)]]
    agg = MetricAggregate(vals, Stats._from_c(st)) if vals else None
    return MetricsResults(n.value, agg)
