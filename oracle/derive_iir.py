"""Charalampidis (2016) truncated-cosine recursive Gaussian, sigma = 1.5.

Test infrastructure (part of the oracle): solves the IIR coefficients in double
precision and rounds them to f32, restating
crates/ssimulacra2-cuda-kernel/build.rs:28-140 (equation numbers below are the
paper's). The f32 values must equal the constants cpu.rs uses
(crates/ssimulacra2-cuda/examples/cpu.rs:931-948); tests/test_oracle_tables.py
checks that, and checks that the double-precision IIR collapses to a symmetric
9-tap FIR (the blur's closed form).

Run as a script to print the constants as C hex literals.
"""
import math

import numpy as np

SIGMA = 1.5


def solve(sigma=SIGMA):
    """Return a dict of the double-precision coefficients."""
    radius = round(3.2795 * sigma + 0.2546)          # (57)
    base = math.pi / (2.0 * radius)                   # Table I
    omega = np.array([base, 3.0 * base, 5.0 * base])
    p = np.array([1.0 / math.tan(0.5 * omega[0]),     # (37)
                  -1.0 / math.tan(0.5 * omega[1]),
                  1.0 / math.tan(0.5 * omega[2])])
    r = np.array([p[0] * p[0] / math.sin(omega[0]),   # (44)
                  -p[1] * p[1] / math.sin(omega[1]),
                  p[2] * p[2] / math.sin(omega[2])])
    rho = np.exp(-0.5 * sigma * sigma * omega * omega) / radius   # (50)
    d13 = p[0] * r[1] - r[0] * p[1]                   # (52)
    d35 = p[1] * r[2] - r[1] * p[2]
    d51 = p[2] * r[0] - r[2] * p[0]
    zeta15 = d35 / d13
    zeta35 = d51 / d13
    a = np.array([[p[0], p[1], p[2]],                 # (56)
                  [r[0], r[1], r[2]],
                  [zeta15, zeta35, 1.0]])
    gamma = np.array([1.0, radius * radius - sigma * sigma,          # (55)
                      zeta15 * rho[0] + zeta35 * rho[1] + rho[2]])
    beta = np.linalg.solve(a, gamma)                  # (53)
    assert abs(float(beta @ p) - 1.0) < 1e-12        # (39): unit DC gain
    n2 = -beta * np.cos(omega * (radius + 1.0))      # (33)
    d1 = -2.0 * np.cos(omega)
    return {"radius": radius, "n2": n2, "d1": d1}


def f32_constants(sigma=SIGMA):
    """The f32 constants in the two recurrence forms cpu.rs uses.

    Row form    o = MUL_PREV*prev + (MUL_PREV2*prev2 + MUL_IN*sum)
    Column form o = VERT_MUL_IN*sum - (VERT_MUL_PREV*prev + prev2)
    """
    s = solve(sigma)
    out = {"RADIUS": s["radius"]}
    for i, k in enumerate((1, 3, 5)):
        out[f"MUL_IN_{k}"] = np.float32(s["n2"][i])
        out[f"MUL_PREV_{k}"] = np.float32(-s["d1"][i])
        out[f"MUL_PREV2_{k}"] = np.float32(-1.0)
        out[f"VERT_MUL_IN_{k}"] = np.float32(s["n2"][i])
        out[f"VERT_MUL_PREV_{k}"] = np.float32(s["d1"][i])
    return out


def iir_row_f64(x):
    """Double-precision row-form IIR with zero padding (cpu.rs:967-1022 order)."""
    s = solve()
    n_big = s["radius"]
    w = len(x)
    prev = np.zeros(3)
    prev2 = np.zeros(3)
    out = np.zeros(w)
    for n in range(-n_big + 1, w):
        left, right = n - n_big - 1, n + n_big - 1
        sm = (x[left] if left >= 0 else 0.0) + (x[right] if right < w else 0.0)
        o = s["n2"] * sm - s["d1"] * prev - prev2
        prev2 = prev
        prev = o
        if n >= 0:
            out[n] = o.sum()
    return out


def fir_taps_f64():
    """Centre-first taps h[0..4] of the FIR the IIR equals (impulse response)."""
    w = 41
    delta = np.zeros(w)
    delta[20] = 1.0
    resp = iir_row_f64(delta)
    return resp[20:25].copy(), resp


if __name__ == "__main__":
    c = f32_constants()
    for k, v in c.items():
        if k == "RADIUS":
            print(f"#define RADIUS {v}")
        else:
            print(f"{k} = {float(v).hex()}f  /* {float(v)!r} */")
    taps, _ = fir_taps_f64()
    print("FIR taps", [float(t).hex() for t in taps])
