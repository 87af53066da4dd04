"""Writes the small fixtures under tests/golden/.

splitmix64_kat.json: known answers of the published splitmix64 finaliser
(Steele, Lea & Flood 2014; the constants 0x9e3779b97f4a7c15,
0xbf58476d1ce4e5b9, 0x94d049bb133111eb), computed here with Python's
arbitrary-precision integers -- independent of the C oracle it checks.

noise_kat.json: values of the counter-based white noise of
`noise_kernel!` (src/spectral/ic.jl:38-48) at chosen (n, seed, i, j, k, c)
(0-based grid indices, component c):
  key = ((c·n + k)·n + j)·n + i,
  r1 = splitmix64(seed + 2 key), r2 = splitmix64(seed + 2 key + 1)  (mod 2^64),
  u = ((r >> 11) + 0.5)·2^-53   (IEEE double, as Julia's Float64 promotion),
  v = sqrt(−2 ln u1)·cospi(2 u2)  (Box–Muller),
with cospi evaluated by exact reduction to [0, 1/4] before one libm call, so
each value is within a few ulp of the exact result.
"""
import json
import math
import os

M = (1 << 64) - 1


def splitmix64(x):
    z = (x + 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def cospi(x):
    x = math.fmod(abs(x), 2.0)
    if x > 1.0:
        x = 2.0 - x  # exact
    if x > 0.5:
        return -cospi(1.0 - x)
    if x > 0.25:
        return math.sin(math.pi * (0.5 - x))
    return math.cos(math.pi * x)


def noise_value(n, seed, i, j, k, c):
    key = ((c * n + k) * n + j) * n + i
    r1 = splitmix64((seed + 2 * key) & M)
    r2 = splitmix64((seed + 2 * key + 1) & M)
    u1 = (float(r1 >> 11) + 0.5) * 2.0 ** -53
    u2 = (float(r2 >> 11) + 0.5) * 2.0 ** -53
    return math.sqrt(-2.0 * math.log(u1)) * cospi(2.0 * u2)


NOISE_POINTS = [
    (12, 0, 0, 0, 0, 0), (12, 0, 1, 0, 0, 0), (12, 0, 0, 1, 0, 0), (12, 0, 0, 0, 1, 0), (12, 0, 0, 0, 0, 1),
    (12, 0, 11, 11, 11, 2), (12, 0, 3, 7, 5, 1), (12, 0, 5, 3, 7, 1), (12, 0, 7, 5, 3, 2),
    (12, 7, 2, 9, 4, 0), (12, 7, 4, 9, 2, 2), (12, 2 ** 40 + 5, 6, 0, 11, 1),
    (18, 0, 17, 0, 0, 0), (18, 0, 0, 17, 0, 2), (18, 0, 0, 0, 17, 1), (18, 0, 9, 4, 13, 2),
    (18, 3, 13, 4, 9, 0), (18, 2 ** 63 + 1, 1, 2, 3, 1), (18, M, 16, 15, 14, 2),
]


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    xs = [0, 1, 2, 7, 14, 15, 12345, 2 ** 32, 2 ** 63, M]
    with open(os.path.join(here, "splitmix64_kat.json"), "w") as f:
        json.dump({"pairs": [[str(x), f"{splitmix64(x):#018x}"] for x in xs]}, f, indent=1)
    with open(os.path.join(here, "noise_kat.json"), "w") as f:
        json.dump({"fields": ["n", "seed", "i", "j", "k", "c", "value"],
                   "points": [[n, str(seed), i, j, k, c, repr(noise_value(n, seed, i, j, k, c))]
                              for (n, seed, i, j, k, c) in NOISE_POINTS]}, f, indent=1)
