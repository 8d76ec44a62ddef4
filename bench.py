#!/usr/bin/env python3
"""ssimulacra2 throughput on one MI355X.

Headline workload (`value`): one 1920x1080 8-bit sRGB pair per step through
the exact f32 path (value noise with 5 hard-edged rectangles vs its 8x8-DCT
quantised copy, step 16 / DC step 8, seed 2), inputs already resident in HBM.
A step is one full ssimulacra2 evaluation: pyramid + XYB, row IIR, column IIR
with the error maps and f64 sums, on-device finalisation, 8-byte score copy.

The same JSON line also carries
  batched_fast  256 uneven pairs (w, h log-uniform in [64, 2048]) per launch of
                the fast form, FP16 and FP32 storage, autotuned tiles;
  serving       Poisson requests of the mixed-size stream through the
                continuous-batching server (FP16 storage), p50/p99 latency;
  roofline      the longest-running exact-path kernel, timed with HIP events
                on its own stream, against the 8 TB/s HBM peak;
  cpu_baseline  the oracle's cpu.rs restatement timed on this host's cores on
                a bounded sample; the same leg checks the fast-form scores
                against the FP64 oracle (accuracy columns).

  video         3840x2160 10-bit BT.709 4:2:0 frame pairs (value noise drifting
                2 px/frame vs 6-bit luma banding + 3x3 chroma blur), converted
                on the GPU inside the first kernel; exact f32 path per frame
                and the batched fast form with f32/f16 storage (bfloat16 is
                refused by the library), |dscore| against the FP64 oracle per
                frame. By default 24 of the 240
                frames (every 10th); `--video-frames 240` runs all of them.

`--workload serving --serving-seconds 60` runs the full serving workload alone.
"""
import argparse
import ctypes
import json
import os
import statistics
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
PKG_DIR = ROOT / "turbo-metrics_amd"
sys.path.insert(0, str(PKG_DIR))

METRIC = "ssimulacra2 Mpx/s at 1080p + mixed-size p99 latency at |Δscore|≤0.05"
HBM_PEAK_GBS = 8000.0
W, H, SEED, STEP = 1920, 1080, 2, 16.0


# Largest batch of the serving and mixed-stream legs: 1024 pairs, 128 Mpx
# (about 2.2 GB of binary16 planes per engine). Measured on the mixed-size
# stream with two engines: 64 Mpx batches 43.0, 128 Mpx 45.1, 256 Mpx 45.3
# Gpx/s; larger launches have fewer ramp and tail phases.
BATCH_PAIRS, BATCH_PIXELS = 1024, 128 << 20


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def scale_dims(w, h):
    dims = []
    for s in range(6):
        if w < 8 or h < 8:
            break
        if s > 0:
            w, h = (w + 1) // 2, (h + 1) // 2
        dims.append((w, h))
    return dims


def decomposition_bytes(w, h, S=4, I=6):
    """HBM bytes per pair of the unfused reference decomposition (S = bytes per
    stored plane sample, I = input bytes per scale-0 pixel for the pair):
      ingest (scale 0) I + 12S, downscale 7.5S per scale that has a next one,
      xyb (scales >= 1) 12S, blur A 21S, blur B 30S, maps/partials 21S
    per pixel of each scale. The fused kernels never write most of these
    planes, so this prices the pair, not a kernel."""
    n = [a * b for a, b in scale_dims(w, h)]
    return int((I + 12 * S) * n[0] + 7.5 * S * sum(n[:-1]) + 12 * S * sum(n[1:]) + 72 * S * sum(n))


def compulsory_bytes(w, h, S=4, I=6):
    """Bytes each fused exact-path kernel must move per pair (n_s = pixels of
    scale s):
      k_pyramid  the input, I per scale-0 pixel, and the six XYB planes it
                 stores, 6S per pixel of every scale;
      k_rows     the six XYB planes read, 6S, and the fifteen row-pass planes
                 written (5 quantities x 3 channels), 15S;
      k_cols     the fifteen row-pass planes and the six XYB planes read, 21S,
                 plus 48 B of partial sums per (scale, channel, 32-column strip);
      finalize   the partial sums read and the 121-double result record."""
    dims = scale_dims(w, h)
    n = [a * b for a, b in dims]
    parts = sum(3 * ((a + 31) // 32) for a, _ in dims) * 48
    return {"pyramid": I * n[0] + 6 * S * sum(n), "rows": 21 * S * sum(n),
            "cols": 21 * S * sum(n) + parts, "finalize": parts + 121 * 8}


def fast_compulsory_bytes(spec, S):
    """Bytes the batched fast form must move per batch: each pair's input
    (6 B per scale-0 pixel), its six XYB planes of every scale stored by the
    pyramid (6S per pixel) and read back once by the tile kernel (6S)."""
    total = 0
    for _, _, w, h, _ in spec:
        n = [a * b for a, b in scale_dims(w, h)]
        total += 6 * n[0] + 12 * S * sum(n)
    return total


def load_traffic(kernel):
    """HBM bytes per launch from the committed rocprofv3 counter summary, if any."""
    p = ROOT / "profiles" / "pmc_1080p_exact_f32.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text()).get("traffic_bytes", {}).get(kernel)
    except Exception:
        return None


# ---------------------------------------------------------------- GPU legs

def run_plain(s2, synth, steps, warmup):
    ref, dis = synth.pair(synth.NOISE_RECT_DCT, SEED, W, H, STEP)
    m = s2.Ssimulacra2(W, H)
    dr, dd = s2.DeviceImage.from_host(ref), s2.DeviceImage.from_host(dis)
    stream = s2.Stream()
    for _ in range(warmup):
        m.compute_srgb(dr, dd, stream)
    score = m.get_score()
    s2.device_synchronize()
    stream.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        m.compute_srgb(dr, dd, stream)
    stream.sync()
    dt = time.perf_counter() - t0
    assert m.get_score() == score
    lat = []
    for _ in range(min(steps, 100)):
        t = time.perf_counter()
        m.compute_srgb(dr, dd, stream)
        m.get_score()
        lat.append((time.perf_counter() - t) * 1e3)
    lat.sort()
    kms = m.time_kernels(dr, dd, reps=max(5, min(steps, 50)), stream=stream)
    # host bytes in (compute_from_cpu_srgb_sync): H2D of both images + compute + sync
    n_host = max(5, min(steps, 50))
    assert m.compute_from_cpu_srgb_sync(ref, dis, stream) == score
    t0 = time.perf_counter()
    for _ in range(n_host):
        m.compute_from_cpu_srgb_sync(ref, dis, stream)
    host_dt = (time.perf_counter() - t0) / n_host
    m.close()
    return {"ref": ref, "dis": dis, "score": score, "dt": dt, "lat": lat, "kms": kms, "host_dt": host_dt}


def run_batch(s2, workloads, storage, host_pairs, spec, reps):
    dev = [(s2.DeviceImage.from_host(r), s2.DeviceImage.from_host(d)) for r, d in host_pairs]
    px = workloads.total_pixels(spec)
    b = s2.Ssim2Batch(max_pairs=len(dev), max_pixels=px + 1024, storage=storage)
    stream = s2.Stream()
    tune = b.autotune(dev, reps=2, stream=stream)
    scores = b.compute(dev, stream)  # warm (graph capture of the first staging slot)
    for _ in range(3):  # untimed: the second slot's graph, clocks
        b.compute_async(dev, stream)
    stream.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.compute_async(dev, stream)
    stream.sync()
    dt = (time.perf_counter() - t0) / reps
    again = b.get_scores()
    assert np.array_equal(again, scores)
    cfg, mult = b.tiling()
    # the same pairs with every reference pyramid resident (registered once,
    # untimed): the rate-distortion loop's case of one reference, many
    # distorted images
    ids = [b.add_reference(r, stream) for r, _ in dev]
    rpairs = [(i, d) for i, (_, d) in zip(ids, dev)]
    for _ in range(3):
        b.compute_ref_async(rpairs, stream)
    stream.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.compute_ref_async(rpairs, stream)
    stream.sync()
    dt_ref = (time.perf_counter() - t0) / reps
    assert np.array_equal(b.get_scores(), scores)
    b.close()
    sbytes = 4 if storage == "f32" else 2
    comp = fast_compulsory_bytes(spec, sbytes)
    return {
        "storage": storage,
        "pairs": len(dev),
        "roofline": {"compulsory_bytes": int(comp), "achieved_gbs": round(comp / dt / 1e9, 1),
                     "frac": round(comp / dt / 1e9 / HBM_PEAK_GBS, 4),
                     "model": "input + XYB planes stored once and read once per batch"},
        "mpx_per_batch": round(px / 1e6, 3),
        "ms_per_batch": round(dt * 1e3, 4),
        "value": round(px / 1e6 / dt, 1),
        "unit": "Mpx/s",
        "tile": list(s2.TILE_SHAPES[cfg]),
        "tile_kernel": "matrix cores (binary16 hi/lo operands)" if (cfg == 4 and storage == "f16") else "VALU FIR",
        "grid_x_cu": mult,
        "resident_ref": {"ms_per_batch": round(dt_ref * 1e3, 4), "value": round(px / 1e6 / dt_ref, 1),
                         "unit": "Mpx/s", "note": "reference pyramids resident, same scores"},
        "autotune": tune.strip().splitlines(),
        "scores": scores,
    }


def run_exact_pairs(s2, host_pairs, spec):
    """The plain exact f32 path's score of every pair (one context per pair)."""
    out = []
    for (r, d), (_, _, w, h, _) in zip(host_pairs, spec):
        with s2.Ssimulacra2(w, h) as m:
            out.append(m.compute_from_cpu_srgb_sync(r, d))
    return out


def run_serving(s2, workloads, seconds, rate, storage="f16", slots=3, resident_refs=False):
    """The Poisson stream through the server. With resident_refs every pool
    reference's pyramid is registered once and each request names it
    (ssim2_server_submit_ref) instead of passing the reference image."""
    spec, cls = workloads.serving_pool_spec()
    per_class = [cls.count(c) for c in range(4)]
    host = workloads.generate(spec)
    dev = [(s2.DeviceImage.from_host(r), s2.DeviceImage.from_host(d)) for r, d in host]
    px_of = [w * h for _, _, w, h, _ in spec]
    t_arr, pick = workloads.serving_schedule(rate, seconds, pool_per_class=per_class)
    sv = s2.Ssim2Server(storage=storage, slots=slots, max_pairs_per_batch=BATCH_PAIRS,
                        max_pixels_per_batch=BATCH_PIXELS, hbm_budget=64 << 30)
    # warm-up: every pool pair once (graph capture, caches)
    for i, p in enumerate(dev):
        sv.submit(*p, 10_000_000 + i)
    sv.drain(120.0)
    sv.poll()
    ref_ids = [sv.add_reference(r) for r, _ in dev] if resident_refs else None
    results = []
    t0 = time.perf_counter()
    for i, (ta, k) in enumerate(zip(t_arr, pick)):
        while True:
            now = time.perf_counter() - t0
            if now >= ta:
                break
            if ta - now > 0.002:
                time.sleep(0.001)
        if ref_ids:
            sv.submit_ref(ref_ids[int(k)], dev[int(k)][1], i)
        else:
            sv.submit(*dev[int(k)], i)
        if i % 256 == 0:
            results.extend(sv.poll())
    sv.drain(300.0)
    results.extend(sv.poll())
    st = sv.stats()
    sv.close()
    ok = [r for r in results if r[4] == "SSIM2_OK"]
    lat = np.array(sorted((r[3] - r[2]) * 1e3 for r in ok))
    span = max(r[3] for r in ok) - min(r[2] for r in ok)
    px = sum(px_of[int(pick[r[0]])] for r in ok)
    # every request of a pool pair must carry the same score
    by_pair = {}
    for r in ok:
        by_pair.setdefault(int(pick[r[0]]), set()).add(r[1])
    consistent = all(len(v) == 1 for v in by_pair.values())
    return {
        "storage": storage,
        "rate_hz": rate,
        "seconds": seconds,
        "note": f"{seconds:g} s of the Poisson schedule (the full configuration is 60 s: "
                "--workload serving --serving-seconds 60)",
        "requests": len(t_arr),
        "completed": len(ok),
        "batches": st["batches"],
        "value": round(px / 1e6 / span, 1),
        "unit": "Mpx/s",
        "offered_mpx_s": round(sum(px_of[int(k)] for k in pick) / 1e6 / seconds, 1),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3),
                       "max": round(float(lat[-1]), 3)},
        "same_score_per_pair": consistent,
        "pool_scores": {k: next(iter(v)) for k, v in by_pair.items()},
        "pool_spec": spec,
        "pool_host": host,
    }


def run_mixed(s2, workloads, pool_dev, spec, cls, seconds=1.0, rate=2000.0, reps=3, storage="f16"):
    """Throughput on the mixed-size request stream (the first `seconds` of the
    Poisson schedule, all requests present at once): the plain exact f32 path
    pair by pair (one context per size, async on one stream) against the
    batched fast form with up to 1024 pairs / 128 Mpx per launch (two batch
    engines alternating, each batch's scores read after the next is
    launched)."""
    per_class = [cls.count(c) for c in range(4)]
    _, pick = workloads.serving_schedule(rate, seconds, pool_per_class=per_class)
    reqs = [int(k) for k in pick]
    px = sum(spec[k][2] * spec[k][3] for k in reqs)
    stream = s2.Stream()
    ctxs = {}
    for k in set(reqs):
        wh = (spec[k][2], spec[k][3])
        if wh not in ctxs:
            ctxs[wh] = s2.Ssimulacra2(*wh)
    plain_scores = {}
    for k in set(reqs):  # warm-up, one score per pool pair
        c = ctxs[(spec[k][2], spec[k][3])]
        plain_scores[k] = c.compute_srgb_sync(*pool_dev[k], stream=stream)
    t0 = time.perf_counter()
    for _ in range(reps):
        for k in reqs:
            ctxs[(spec[k][2], spec[k][3])].compute_srgb(*pool_dev[k], stream)
    stream.sync()
    dt_plain = (time.perf_counter() - t0) / reps
    for c in ctxs.values():
        c.close()
    # batches as the server forms them: FIFO, at most BATCH_PAIRS pairs and BATCH_PIXELS
    batches, cur, cur_px = [], [], 0
    for k in reqs:
        p = spec[k][2] * spec[k][3]
        if cur and (len(cur) == BATCH_PAIRS or cur_px + p > BATCH_PIXELS):
            batches.append(cur)
            cur, cur_px = [], 0
        cur.append(k)
        cur_px += p
    if cur:
        batches.append(cur)
    # two engines on two streams, alternating: batch k+1 is staged and
    # launched before batch k's scores are read, as a throughput client would
    engines = [s2.Ssim2Batch(max_pairs=BATCH_PAIRS, max_pixels=BATCH_PIXELS, storage=storage) for _ in range(2)]
    streams = [stream, s2.Stream()]
    engines[0].autotune([pool_dev[k] for k in batches[0]], reps=2, stream=stream)
    engines[1].set_tiling(*engines[0].tiling())
    for i, bt in enumerate(batches):
        engines[i % 2].compute([pool_dev[k] for k in bt], streams[i % 2])
    t0 = time.perf_counter()
    for _ in range(reps):
        pending = None
        for i, bt in enumerate(batches):
            e = engines[i % 2]
            e.compute_async([pool_dev[k] for k in bt], streams[i % 2])
            if pending is not None:
                pending.get_scores()
            pending = e
        pending.get_scores()
    dt_batch = (time.perf_counter() - t0) / reps
    for e in engines:
        e.close()
    plain = px / 1e6 / dt_plain
    batched = px / 1e6 / dt_batch
    return {"requests": len(reqs), "mpx": round(px / 1e6, 1), "batches": len(batches),
            "plain_exact_f32": round(plain, 1), f"batched_{storage}": round(batched, 1),
            "unit": "Mpx/s", "ratio": round(batched / plain, 2)}


VIDEO_W, VIDEO_H, VIDEO_SEED, VIDEO_LEN = 3840, 2160, 3, 240


def video_frames(n):
    """Frame indices of an n-frame sample spread over the 240-frame sequence."""
    n = max(1, min(VIDEO_LEN, n))
    return [int(t) for t in np.linspace(0, VIDEO_LEN, n, endpoint=False).round()]


def run_video(s2, synth, nframes, reps=3):
    ts = video_frames(nframes)
    with ThreadPoolExecutor(max(1, min(16, os.cpu_count() or 1))) as ex:
        host = list(ex.map(lambda t: synth.yuv420_pair(VIDEO_SEED, VIDEO_W, VIDEO_H, t), ts))
    dev = [(s2.DeviceYuv.from_planes(*r), s2.DeviceYuv.from_planes(*d)) for r, d in host]
    px = VIDEO_W * VIDEO_H * len(dev)
    stream = s2.Stream()
    m = s2.Ssimulacra2(VIDEO_W, VIDEO_H)
    exact = []
    for r, d in dev:  # warm-up and scores
        exact.append(m.compute_yuv420_sync(r, d, stream))
    t0 = time.perf_counter()
    for _ in range(reps):
        for r, d in dev:
            m.compute_yuv420(r, d, stream)
    stream.sync()
    dt_exact = (time.perf_counter() - t0) / reps
    m.close()
    legs = {"exact_f32": {"value": round(px / 1e6 / dt_exact, 1), "unit": "Mpx/s",
                          "ms_per_frame": round(dt_exact / len(dev) * 1e3, 4), "scores": exact}}
    for storage in ("f32", "f16"):
        b = s2.Ssim2Batch(max_pairs=len(dev), max_pixels=px + 1024, storage=storage)
        scores = b.compute_yuv(dev, stream)
        stream.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            b.compute_yuv_async(dev, stream)
        stream.sync()
        dt = (time.perf_counter() - t0) / reps
        assert np.array_equal(b.get_scores(), scores)
        b.close()
        legs[f"batched_{storage}"] = {"value": round(px / 1e6 / dt, 1), "unit": "Mpx/s",
                                      "ms_per_frame": round(dt / len(dev) * 1e3, 4), "scores": list(scores)}
    del dev
    return {"frames": ts, "width": VIDEO_W, "height": VIDEO_H, "format": "yuv420p10", "legs": legs, "host": host}


def video_accuracy(video):
    """|dscore| of every video leg against the FP64 oracle, per frame (checker only)."""
    oracle = ctypes.CDLL(str(ROOT / "oracle" / "libssim2_oracle.so"))
    vp, i = ctypes.c_void_p, ctypes.c_int
    oracle.oracle_score_yuv420p10_f64.argtypes = [vp] * 6 + [i, i, ctypes.POINTER(ctypes.c_double)]

    def f64(pair):
        planes = [np.ascontiguousarray(p) for p in (*pair[0], *pair[1])]
        s = ctypes.c_double()
        oracle.oracle_score_yuv420p10_f64(*[p.ctypes.data for p in planes], VIDEO_W, VIDEO_H, ctypes.byref(s))
        return s.value

    # about 2 GB of host memory per 4K FP64 evaluation
    with ThreadPoolExecutor(max(1, min(16, os.cpu_count() or 1))) as ex:
        ref = list(ex.map(f64, video["host"]))
    video["f64_scores"] = [round(v, 6) for v in ref]
    for leg in video["legs"].values():
        d = [abs(a - b) for a, b in zip(leg["scores"], ref)]
        leg["abs_delta_vs_f64"] = [round(x, 5) for x in d]
        leg["max_abs_delta_vs_f64"] = round(max(d), 5)
        leg["scores"] = [round(x, 6) for x in leg["scores"]]


# ---------------------------------------------------------------- CPU leg (oracle)

def host_cpu():
    """(cores this process may run on, CPU model from /proc/cpuinfo)."""
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return allowed, model


def cpu_leg(plain, batch_legs, serving, batch_host, batch_spec, config1=None, batch_exact=None):
    """The oracle: CPU baseline timing and the accuracy columns."""
    oracle = ctypes.CDLL(str(ROOT / "oracle" / "libssim2_oracle.so"))
    vp, i = ctypes.c_void_p, ctypes.c_int
    oracle.oracle_score_srgb_many.argtypes = [vp, vp, vp, vp, i, i, i, vp]
    oracle.oracle_score_srgb.argtypes = [vp, vp, i, i, i, ctypes.POINTER(ctypes.c_double), vp, vp]
    # every core this process may use, up to the 16-core host share a
    # one-GPU box grants (its other cores belong to other jobs)
    allowed, model = host_cpu()
    threads = max(1, min(16, allowed))

    def score(ref, dis, mode):
        s = ctypes.c_double()
        r, d = np.ascontiguousarray(ref), np.ascontiguousarray(dis)
        oracle.oracle_score_srgb(r.ctypes.data, d.ctypes.data, r.shape[1], r.shape[0], mode, ctypes.byref(s),
                                 None, None)
        return s.value

    ref, dis = plain["ref"], plain["dis"]
    t0 = time.perf_counter()
    cpu_score = score(ref, dis, 0)
    single = time.perf_counter() - t0
    npairs = 2 * threads
    refs = (ctypes.c_void_p * npairs)(*([ref.ctypes.data] * npairs))
    diss = (ctypes.c_void_p * npairs)(*([dis.ctypes.data] * npairs))
    ws = (ctypes.c_int * npairs)(*([W] * npairs))
    hs = (ctypes.c_int * npairs)(*([H] * npairs))
    out = np.zeros(npairs)
    t0 = time.perf_counter()
    oracle.oracle_score_srgb_many(refs, diss, ws, hs, npairs, 0, threads, out.ctypes.data)
    dt = time.perf_counter() - t0
    cb = {
        "value": round(npairs * W * H / 1e6 / dt, 3),
        "unit": "Mpx/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{npairs} evaluations of the 1920x1080 pair on {threads} std::threads (oracle f32 "
                  f"restatement of cpu.rs, built -O3 -march=x86-64-v3, {dt:.1f} s wall); single-thread "
                  f"{single * 1e3:.0f} ms/pair",
        "single_thread_ms_per_pair": round(single * 1e3, 1),
        "nproc": os.cpu_count(),
        "cores_allowed": allowed,
        "cpu_model": model,
    }
    if config1 is not None:
        # the 512x512 CPU-path configuration: the oracle's score and time,
        # beside the exact GPU path's score of the same pair
        r1, d1, gpu1 = config1
        t0 = time.perf_counter()
        s1 = score(r1, d1, 0)
        t1 = time.perf_counter() - t0
        cb["config_512x512"] = {"cpu_score": s1, "cpu_ms_single_thread": round(t1 * 1e3, 1),
                                "gpu_exact_score": gpu1, "abs_delta": abs(s1 - gpu1),
                                "pair": "4 octaves of value noise + gradient vs Gaussian noise sigma 4 and 8x8 "
                                        "block means on 10% of blocks, seed 1"}
    acc = {"exact_vs_cpu_mirror": abs(cpu_score - plain["score"]),
           "exact_vs_f64": None}
    with ThreadPoolExecutor(threads) as ex:
        acc["exact_vs_f64"] = abs(ex.submit(score, ref, dis, 1).result() - plain["score"])
        # fast form vs FP64 on a stratified sample of the uneven batch (every 8th pair)
        if batch_legs:
            idx = list(range(0, len(batch_spec), 8))
            f64 = list(ex.map(lambda k: score(batch_host[k][0], batch_host[k][1], 1), idx))
            for leg in batch_legs:
                d = [abs(leg["scores"][k] - f) for k, f in zip(idx, f64)]
                leg["max_abs_delta_vs_f64"] = round(max(d), 5)
                leg["delta_sample"] = f"{len(idx)} of {len(batch_spec)} pairs (every 8th)"
            if batch_exact is not None:
                acc["uneven_exact_path_vs_f64_max"] = round(max(abs(batch_exact[k] - f) for k, f in zip(idx, f64)), 5)
    if serving:
        pool_accuracy(serving)
    return cb, acc


def pool_accuracy(serving):
    """Served scores of the pool pairs against the FP64 oracle (checker only)."""
    oracle = ctypes.CDLL(str(ROOT / "oracle" / "libssim2_oracle.so"))
    vp, i = ctypes.c_void_p, ctypes.c_int
    oracle.oracle_score_srgb.argtypes = [vp, vp, i, i, i, ctypes.POINTER(ctypes.c_double), vp, vp]

    def f64(k):
        r, d = serving["pool_host"][k]
        s = ctypes.c_double()
        oracle.oracle_score_srgb(r.ctypes.data, d.ctypes.data, r.shape[1], r.shape[0], 1, ctypes.byref(s), None, None)
        return s.value

    keys = sorted(serving["pool_scores"])
    with ThreadPoolExecutor(max(1, min(16, os.cpu_count() or 1))) as ex:
        ref = list(ex.map(f64, keys))
    serving["max_abs_delta_vs_f64"] = round(max(abs(serving["pool_scores"][k] - v) for k, v in zip(keys, ref)), 5)


# ---------------------------------------------------------------- main

def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", choices=["all", "plain", "batch", "serving", "video", "mixed"], default="all")
    ap.add_argument("--serving-seconds", type=float, default=8.0)
    ap.add_argument("--serving-rate", type=float, default=2000.0)
    ap.add_argument("--batch-reps", type=int, default=5)
    ap.add_argument("--video-frames", type=int, default=24,
                    help="frames of the 240-frame 4K YUV sequence to run (0 skips the video leg)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="also run the plain copy kernel once over this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); pairs are independent and the "
              "multi-GPU form is not part of this build.", file=sys.stderr)
        sys.exit(2)

    import ssimulacra2_hip as s2
    from ssimulacra2_hip import synth, workloads

    if args.workload == "serving":
        log(f"serving {args.serving_seconds} s at {args.serving_rate} req/s")
        sv = run_serving(s2, workloads, args.serving_seconds, args.serving_rate)
        if not args.no_cpu_baseline:
            log("accuracy of the served pool against the FP64 oracle")
            pool_accuracy(sv)
        out = {k: v for k, v in sv.items() if not k.startswith("pool_")}
        out.update({"metric": "ssimulacra2 serving p99 latency (mixed sizes, Poisson)", "n_gpus": 1,
                    "higher_is_better": True, "data": "synthetic"})
        print(json.dumps(out), flush=True)
        return

    if args.workload == "batch":
        bspec = workloads.uneven_batch_spec()
        bhost = workloads.generate(bspec)
        legs = [run_batch(s2, workloads, st, bhost, bspec, args.batch_reps) for st in ("f16", "f32")]
        print(json.dumps({"metric": "ssimulacra2 Mpx/s, 256 uneven pairs per launch", "n_gpus": 1,
                          "legs": [{k: v for k, v in leg.items() if k != "scores"} for leg in legs]}), flush=True)
        return

    if args.workload == "mixed":
        spec, cls = workloads.serving_pool_spec()
        pool_dev = [(s2.DeviceImage.from_host(r), s2.DeviceImage.from_host(d)) for r, d in workloads.generate(spec)]
        mixed = run_mixed(s2, workloads, pool_dev, spec, cls)
        print(json.dumps({"metric": "ssimulacra2 Mpx/s on the mixed-size stream", "n_gpus": 1, **mixed}), flush=True)
        return

    if args.workload == "video":
        log(f"video: {args.video_frames} frames of 3840x2160 yuv420p10")
        video = run_video(s2, synth, args.video_frames)
        if not args.no_cpu_baseline:
            log("video accuracy against the FP64 oracle")
            video_accuracy(video)
        out = {k: v for k, v in video.items() if k != "host"}
        out.update({"metric": "ssimulacra2 Mpx/s on 4K 10-bit YUV 4:2:0 frames", "n_gpus": 1,
                    "higher_is_better": True, "data": "synthetic"})
        print(json.dumps(out), flush=True)
        return

    log("plain 1920x1080 exact path")
    plain = run_plain(s2, synth, args.steps, args.warmup)
    comp = compulsory_bytes(W, H)
    kms = plain["kms"]
    dom = max(("pyramid", "rows", "cols"), key=lambda k: kms[k])
    achieved = comp[dom] / (kms[dom] * 1e-3) / 1e9
    per_kernel = {k: {"ms": round(kms[k], 5), "compulsory_bytes": comp[k],
                      "achieved_gbs": round(comp[k] / (kms[k] * 1e-3) / 1e9, 1),
                      "frac": round(comp[k] / (kms[k] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                      "counted_bytes": load_traffic(k)} for k in ("pyramid", "rows", "cols", "finalize")}
    pair_comp = sum(comp.values())
    pair_s = plain["dt"] / args.steps

    batch_legs, serving, bhost, bspec, mixed, batch_exact = [], None, None, None, None, None
    if args.workload == "all":
        bspec = workloads.uneven_batch_spec()
        log(f"generating the uneven batch ({workloads.total_pixels(bspec) / 1e6:.1f} Mpx)")
        bhost = workloads.generate(bspec)
        for storage in ("f16", "f32"):
            log(f"batched fast form, {storage} storage")
            batch_legs.append(run_batch(s2, workloads, storage, bhost, bspec, args.batch_reps))
        log("the same 256 pairs through the plain exact path, pair by pair")
        exact = run_exact_pairs(s2, bhost, bspec)
        for leg in batch_legs:
            d = np.abs(np.array(leg["scores"]) - np.array(exact))
            leg["vs_plain_exact_path"] = {"max_abs_delta": round(float(d.max()), 5),
                                          "mean_abs_delta": round(float(d.mean()), 5),
                                          "pairs": len(exact),
                                          "note": "the exact path is the reference's f32 arithmetic, itself up "
                                                  "to about 1 from FP64 on these pairs"}
        batch_exact = exact
        log(f"serving {args.serving_seconds} s at {args.serving_rate} req/s")
        serving = run_serving(s2, workloads, args.serving_seconds, args.serving_rate)
        log("serving above capacity (40000 req/s for 1 s)")
        sat = run_serving(s2, workloads, 1.0, 40000.0)
        serving["saturated"] = {"rate_hz": sat["rate_hz"], "seconds": sat["seconds"], "requests": sat["requests"],
                                "completed": sat["completed"], "value": sat["value"], "unit": sat["unit"],
                                "offered_mpx_s": sat["offered_mpx_s"], "latency_ms": sat["latency_ms"],
                                "same_score_per_pair": sat["same_score_per_pair"]}
        log("serving with resident reference pyramids")
        sref = run_serving(s2, workloads, min(args.serving_seconds, 4.0), args.serving_rate, resident_refs=True)
        serving["resident_refs"] = {
            "seconds": sref["seconds"], "requests": sref["requests"], "completed": sref["completed"],
            "value": sref["value"], "unit": sref["unit"], "latency_ms": sref["latency_ms"],
            "same_score_per_pair": sref["same_score_per_pair"],
            "equals_direct": all(serving["pool_scores"].get(k, v) == v for k, v in sref["pool_scores"].items())}
        log("mixed-size stream: plain exact path vs batched fast form")
        pool_dev = [(s2.DeviceImage.from_host(r), s2.DeviceImage.from_host(d)) for r, d in serving["pool_host"]]
        spec, cls = workloads.serving_pool_spec()
        mixed = run_mixed(s2, workloads, pool_dev, spec, cls)
        del pool_dev
    video = None
    if args.workload == "all" and args.video_frames > 0:
        log(f"video: {args.video_frames} frames of 3840x2160 yuv420p10")
        video = run_video(s2, synth, args.video_frames)

    if args.calibrate_copy:
        nbytes = args.calibrate_copy * (1 << 20)
        a, b = s2.DeviceBuffer(nbytes), s2.DeviceBuffer(nbytes)
        st = s2.Stream()
        b.copy_from(a, nbytes, st)
        st.sync()

    mpx = W * H / 1e6
    out = {
        "metric": METRIC,
        "value": round(args.steps * mpx / plain["dt"], 3),
        "unit": "Mpx/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(plain["dt"] / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "srgb_u8_pair_1920x1080_exact_f32",
            "width": W,
            "height": H,
            "pairs_per_step": 1,
            "content": "value noise (256/64/16/4 px) + 5 rectangles vs 8x8 DCT quantised (step 16, DC 8)",
            "seed": SEED,
        },
        "score": plain["score"],
        "latency_ms": {"p50": round(statistics.median(plain["lat"]), 4),
                       "p99": round(plain["lat"][int(0.99 * (len(plain["lat"]) - 1))], 4)},
        "kernels_ms": {k: round(v, 5) for k, v in kms.items()},
        "pcie_inclusive": {"value": round(mpx / plain["host_dt"], 1), "unit": "Mpx/s",
                           "ms_per_pair": round(plain["host_dt"] * 1e3, 4),
                           "note": "host u8 bytes in, H2D + compute + sync per pair (not the headline)"},
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": load_traffic(dom),
            "compulsory_bytes": comp[dom],
            "model": "bytes the fused kernel must move: planes it reads and stores, its input and partial sums",
            "kernels": per_kernel,
            "pair": {"compulsory_bytes": pair_comp,
                     "achieved_gbs": round(pair_comp / pair_s / 1e9, 1),
                     "frac": round(pair_comp / pair_s / 1e9 / HBM_PEAK_GBS, 4),
                     "decomposition_bytes": decomposition_bytes(W, H),
                     "decomposition_note": "the unfused reference decomposition's bytes per pair, for comparison "
                                           "with the reference's layout; the fused path does not move them"},
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline and accuracy columns (oracle)")
        r1, d1 = synth.pair(synth.NOISE_GRAD_BLOCKMEAN, 1, 512, 512)
        with s2.Ssimulacra2(512, 512) as m1:
            gpu1 = m1.compute_from_cpu_srgb_sync(r1, d1)
        cb, acc = cpu_leg(plain, batch_legs, serving, bhost, bspec, config1=(r1, d1, gpu1), batch_exact=batch_exact)
        out["cpu_baseline"] = cb
        out["accuracy"] = acc
        if video:
            video_accuracy(video)
    if batch_legs:
        out["batched_fast"] = [{k: v for k, v in leg.items() if k != "scores"} for leg in batch_legs]
    if serving:
        out["serving"] = {k: v for k, v in serving.items() if not k.startswith("pool_")}
    if mixed:
        out["mixed_stream"] = mixed
    if video:
        out["video"] = {k: v for k, v in video.items() if k != "host"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
