"""Write tests/golden/trig_cases.npz: inputs of the refinement's float
atan2, cos and sin (fast/decode.c:740-744) with their correctly rounded
results, computed here with mpmath at 200 bits (independent of any libm).

    python tests/golden/make_trig_golden.py

Inputs:
  hard    the inputs on which rounding a double result to float is hardest
          (the double result within 2^-50 of its magnitude from a float
          rounding boundary): every float in (0, pi/2] for cos and sin
          and their negatives, and those among 2e9 seeded float pairs for
          atan2, found by find_trig_hard.c (built here with gcc -fopenmp)
  random  seeded: atan2 of float pairs of magnitudes 2^-12..2^12 and of the
          refinement's own shape (both signs, one argument tiny), cos and
          sin of floats in [-pi/2, pi/2]
Arrays: <set>_y, <set>_x, <set>_atan2 (float32); <set>_th, <set>_cos,
<set>_sin (float32).
"""
import subprocess
import sys
import tempfile
from pathlib import Path

import mpmath as mp
import numpy as np

HERE = Path(__file__).resolve().parent
mp.mp.prec = 200


def cr_float(v):
    """The float32 nearest to the mpmath value v (ties to even)."""
    f = np.float32(float(v))  # the double nearest v, then to float: may double-round
    # fix by comparing against the neighbouring floats exactly
    cands = [f, np.nextafter(f, np.float32(np.inf)), np.nextafter(f, np.float32(-np.inf))]
    best = min(cands, key=lambda c: (abs(mp.mpf(float(c)) - v), int(np.frombuffer(np.float32(c).tobytes(), np.uint32)[0]) & 1))
    return np.float32(best)


def search(exe, *args):
    out = subprocess.run([str(exe), *args], check=True, capture_output=True, text=True).stdout
    return [tuple(float.fromhex(v) for v in line.split()) for line in out.splitlines() if line.strip()]


def main():
    with tempfile.TemporaryDirectory() as td:
        exe = Path(td) / "find_trig_hard"
        subprocess.run(["gcc", "-O2", "-fopenmp", "-o", str(exe), str(HERE / "find_trig_hard.c"), "-lm"], check=True)
        hard_th = sorted({v[0] for v in search(exe, "cos") + search(exe, "sin")})
        hard_yx = search(exe, "atan2", "2000000000", "7")
    hard_th = hard_th + [-v for v in hard_th]
    rng = np.random.default_rng(20261204)
    n = 4000
    mant = rng.uniform(1, 2, (n, 2)) * np.exp2(rng.integers(-12, 13, (n, 2))) * rng.choice([-1, 1], (n, 2))
    ry = list(mant[:, 0].astype(np.float32))
    rx = list(mant[:, 1].astype(np.float32))
    # the refinement's arguments: -2 cxy and cyy - cxx of an edge's points
    # (one of them tiny on near-axis edges)
    sh = rng.uniform(-50, 50, (1000, 2)).astype(np.float32)
    sh[:300, 0] *= np.float32(1e-6)
    ry += list(sh[:, 0])
    rx += list(sh[:, 1])
    rth = list(rng.uniform(-np.pi / 2, np.pi / 2, 5000).astype(np.float32))
    rth += [np.float32(np.pi / 2), np.float32(-np.pi / 2), np.float32(0.0), np.float32(1e-30), np.float32(np.pi / 4)]
    out = {}
    for name, yx, th in (("hard", hard_yx, hard_th), ("random", list(zip(ry, rx)), rth)):
        y = np.array([a for a, _ in yx], np.float32)
        x = np.array([b for _, b in yx], np.float32)
        t = np.array(th, np.float32)
        out[name + "_y"], out[name + "_x"], out[name + "_th"] = y, x, t
        out[name + "_atan2"] = np.array([cr_float(mp.atan2(mp.mpf(float(a)), mp.mpf(float(b)))) for a, b in zip(y, x)],
                                        np.float32)
        out[name + "_cos"] = np.array([cr_float(mp.cos(mp.mpf(float(v)))) for v in t], np.float32)
        out[name + "_sin"] = np.array([cr_float(mp.sin(mp.mpf(float(v)))) for v in t], np.float32)
    np.savez_compressed(HERE / "trig_cases.npz", **out)
    print("hard: %d atan2, %d cos/sin; random: %d atan2, %d cos/sin" % (len(hard_yx), len(hard_th), len(ry), len(rth)))


if __name__ == "__main__":
    main()
