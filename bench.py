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


def algorithmic_bytes(w, h):
    """Compulsory HBM bytes per launch of each exact-path kernel for one pair.

    pyramid: 6 B/px of u8 input at scale 0 + 6 XYB planes (f32) written per scale;
    rows:    6 XYB planes read + 15 row-pass planes written per scale;
    cols:    15 row-pass planes + 6 XYB planes read per scale (sums only out).
    """
    sn = sum(a * b for a, b in scale_dims(w, h))
    return {"pyramid": 6 * w * h + 24 * sn, "rows": 84 * sn, "cols": 84 * sn, "finalize": 0}


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
    m.close()
    return {"ref": ref, "dis": dis, "score": score, "dt": dt, "lat": lat, "kms": kms}


def run_batch(s2, workloads, storage, host_pairs, spec, reps):
    dev = [(s2.DeviceImage.from_host(r), s2.DeviceImage.from_host(d)) for r, d in host_pairs]
    px = workloads.total_pixels(spec)
    b = s2.Ssim2Batch(max_pairs=len(dev), max_pixels=px + 1024, storage=storage)
    stream = s2.Stream()
    tune = b.autotune(dev, reps=2, stream=stream)
    scores = b.compute(dev, stream)  # warm (graph capture)
    stream.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.compute_async(dev, stream)
    stream.sync()
    dt = (time.perf_counter() - t0) / reps
    again = b.get_scores()
    assert np.array_equal(again, scores)
    cfg, mult = b.tiling()
    b.close()
    return {
        "storage": storage,
        "pairs": len(dev),
        "mpx_per_batch": round(px / 1e6, 3),
        "ms_per_batch": round(dt * 1e3, 4),
        "value": round(px / 1e6 / dt, 1),
        "unit": "Mpx/s",
        "tile": list(s2.TILE_SHAPES[cfg]),
        "grid_x_cu": mult,
        "autotune": tune.strip().splitlines(),
        "scores": scores,
    }


def run_serving(s2, workloads, seconds, rate, storage="f16", slots=3):
    spec, cls = workloads.serving_pool_spec()
    per_class = [cls.count(c) for c in range(4)]
    host = workloads.generate(spec)
    dev = [(s2.DeviceImage.from_host(r), s2.DeviceImage.from_host(d)) for r, d in host]
    px_of = [w * h for _, _, w, h, _ in spec]
    t_arr, pick = workloads.serving_schedule(rate, seconds, pool_per_class=per_class)
    sv = s2.Ssim2Server(storage=storage, slots=slots, max_pairs_per_batch=512,
                        max_pixels_per_batch=64 << 20, hbm_budget=64 << 30)
    # warm-up: every pool pair once (graph capture, caches)
    for i, p in enumerate(dev):
        sv.submit(*p, 10_000_000 + i)
    sv.drain(120.0)
    sv.poll()
    results = []
    t0 = time.perf_counter()
    for i, (ta, k) in enumerate(zip(t_arr, pick)):
        while True:
            now = time.perf_counter() - t0
            if now >= ta:
                break
            if ta - now > 0.002:
                time.sleep(0.001)
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


# ---------------------------------------------------------------- CPU leg (oracle)

def cpu_leg(plain, batch_legs, serving, batch_host, batch_spec):
    """The oracle: CPU baseline timing and the accuracy columns."""
    oracle = ctypes.CDLL(str(ROOT / "oracle" / "libssim2_oracle.so"))
    vp, i = ctypes.c_void_p, ctypes.c_int
    oracle.oracle_score_srgb_many.argtypes = [vp, vp, vp, vp, i, i, i, vp]
    oracle.oracle_score_srgb.argtypes = [vp, vp, i, i, i, ctypes.POINTER(ctypes.c_double), vp, vp]
    threads = max(1, min(16, os.cpu_count() or 1))

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
                  f"restatement of cpu.rs, {dt:.1f} s wall); single-thread {single * 1e3:.0f} ms/pair",
        "single_thread_ms_per_pair": round(single * 1e3, 1),
    }
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
    ap.add_argument("--workload", choices=["all", "plain", "serving"], default="all")
    ap.add_argument("--serving-seconds", type=float, default=8.0)
    ap.add_argument("--serving-rate", type=float, default=2000.0)
    ap.add_argument("--batch-reps", type=int, default=5)
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

    log("plain 1920x1080 exact path")
    plain = run_plain(s2, synth, args.steps, args.warmup)
    alg = algorithmic_bytes(W, H)
    kms = plain["kms"]
    dom = max(("pyramid", "rows", "cols"), key=lambda k: kms[k])
    achieved = alg[dom] / (kms[dom] * 1e-3) / 1e9

    batch_legs, serving, bhost, bspec = [], None, None, None
    if args.workload == "all":
        bspec = workloads.uneven_batch_spec()
        log(f"generating the uneven batch ({workloads.total_pixels(bspec) / 1e6:.1f} Mpx)")
        bhost = workloads.generate(bspec)
        for storage in ("f16", "f32"):
            log(f"batched fast form, {storage} storage")
            batch_legs.append(run_batch(s2, workloads, storage, bhost, bspec, args.batch_reps))
        log(f"serving {args.serving_seconds} s at {args.serving_rate} req/s")
        serving = run_serving(s2, workloads, args.serving_seconds, args.serving_rate)

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
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": load_traffic(dom),
            "algorithmic_bytes": alg[dom],
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline and accuracy columns (oracle)")
        cb, acc = cpu_leg(plain, batch_legs, serving, bhost, bspec)
        out["cpu_baseline"] = cb
        out["accuracy"] = acc
    if batch_legs:
        out["batched_fast"] = [{k: v for k, v in leg.items() if k != "scores"} for leg in batch_legs]
    if serving:
        out["serving"] = {k: v for k, v in serving.items() if not k.startswith("pool_")}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
