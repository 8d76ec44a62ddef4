"""Seeded workload definitions for the ssimulacra2 benchmarks and tests.

* uneven_batch: 256 sRGB pairs, width and height independently log-uniform in
  [64, 2048]; reference = value noise + gradient, distortion = 8x8 DCT
  quantisation with a step uniform in [4, 48] (DC step = step / 2).
* serving_mix: request sizes 50% 256x256, 30% 1920x1080, 15% log-uniform
  64..1024 rectangles, 5% 3840x2160, with a seeded quantisation step; Poisson
  arrivals at a given rate.
Generation runs on a thread pool (the C generators release the GIL).
"""
from __future__ import annotations

import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import synth

UNEVEN_SEED = 4
SERVING_SEED = 5


def _threads(n=None):
    return max(1, min(16, n or os.cpu_count() or 1))


def uneven_batch_spec(n=256, seed=UNEVEN_SEED, lo=64, hi=2048):
    """[(kind, seed, w, h, step)] of the uneven batch."""
    spec = []
    for i in range(n):
        w = synth.log_uniform(seed, i, 1, lo, hi)
        h = synth.log_uniform(seed, i, 2, lo, hi)
        step = 4.0 + 44.0 * synth.uniform(seed, i, 3)
        spec.append((synth.NOISE_GRAD_DCT, synth.mix_seed(seed, i), w, h, step))
    return spec


def serving_size(seed, i):
    u = synth.uniform(seed, i, 10)
    if u < 0.50:
        return 256, 256
    if u < 0.80:
        return 1920, 1080
    if u < 0.95:
        return synth.log_uniform(seed, i, 11, 64, 1024), synth.log_uniform(seed, i, 12, 64, 1024)
    return 3840, 2160


def serving_pool_spec(per_class=(16, 8, 16, 2), seed=SERVING_SEED):
    """Distinct pairs kept resident on the device: per class counts for
    (256x256, 1920x1080, small rectangles, 3840x2160)."""
    spec = []
    cls = []
    k = 0
    for c, count in enumerate(per_class):
        for j in range(count):
            s = synth.mix_seed(seed, 1000 * (c + 1) + j)
            if c == 0:
                w, h = 256, 256
            elif c == 1:
                w, h = 1920, 1080
            elif c == 2:
                w, h = synth.log_uniform(seed, 5000 + j, 11, 64, 1024), synth.log_uniform(seed, 5000 + j, 12, 64, 1024)
            else:
                w, h = 3840, 2160
            step = 4.0 + 44.0 * synth.uniform(seed, 7000 + k, 3)
            spec.append((synth.NOISE_GRAD_DCT, s, w, h, step))
            cls.append(c)
            k += 1
    return spec, cls


def serving_schedule(rate_hz=2000.0, duration_s=60.0, seed=SERVING_SEED, pool_classes=None, pool_per_class=None):
    """Poisson arrival times and the pool index of each request."""
    rng = np.random.default_rng(seed)
    n = int(rate_hz * duration_s * 1.2) + 16
    t = np.cumsum(rng.exponential(1.0 / rate_hz, n))
    t = t[t < duration_s]
    u = rng.random(len(t))
    cls = np.where(u < 0.50, 0, np.where(u < 0.80, 1, np.where(u < 0.95, 2, 3)))
    counts = np.array(pool_per_class)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    pick = starts[cls] + rng.integers(0, 1 << 30, len(t)) % counts[cls]
    return t, pick


def generate(spec, threads=None):
    """[(ref, dis)] host arrays for a spec list, generated in parallel."""
    def one(e):
        kind, s, w, h, step = e
        return synth.pair(kind, s, w, h, step)

    with ThreadPoolExecutor(_threads(threads)) as ex:
        return list(ex.map(one, spec))


def total_pixels(spec):
    return sum(w * h for _, _, w, h, _ in spec)
