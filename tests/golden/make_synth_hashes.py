"""Write SHA-256 digests of small synthetic pairs (tests/golden/synth_sha256.json).

The generators (turbo-metrics_amd/csrc/synth.c) must be byte-deterministic on
every host; tests/test_synth.py recomputes these digests. Re-run only when the
generators change on purpose.
"""
import hashlib
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent / "turbo-metrics_amd"))
from ssimulacra2_hip import synth  # noqa: E402

CASES = [(1, 1, 64, 48, 16.0), (2, 2, 96, 80, 16.0), (4, 3, 33, 17, 8.0), (4, 4, 40, 24, 40.0)]


def digests():
    out = {}
    for kind, seed, w, h, step in CASES:
        r, d = synth.pair(kind, seed, w, h, step)
        out[f"pair_{kind}_{seed}_{w}x{h}_{step:g}"] = [hashlib.sha256(r.tobytes()).hexdigest(),
                                                       hashlib.sha256(d.tobytes()).hexdigest()]
    (yr, ur, vr), (yd, ud, vd) = synth.yuv420_pair(5, 64, 36, 3)
    out["yuv420_5_64x36_t3"] = [hashlib.sha256(b"".join(p.tobytes() for p in (yr, ur, vr))).hexdigest(),
                                hashlib.sha256(b"".join(p.tobytes() for p in (yd, ud, vd))).hexdigest()]
    out["log_uniform_7"] = [synth.log_uniform(7, i, 1, 64, 2048) for i in range(8)]
    return out


if __name__ == "__main__":
    (HERE / "synth_sha256.json").write_text(json.dumps(digests(), indent=1) + "\n")
    print("wrote synth_sha256.json")
