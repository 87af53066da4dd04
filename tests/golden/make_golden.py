"""Writes the small fixtures under tests/golden/.

splitmix64_kat.json: known answers of the published splitmix64 finaliser
(Steele, Lea & Flood 2014; the constants 0x9e3779b97f4a7c15,
0xbf58476d1ce4e5b9, 0x94d049bb133111eb), computed here with Python's
arbitrary-precision integers -- independent of the C oracle it checks.
"""
import json
import os

M = (1 << 64) - 1


def splitmix64(x):
    z = (x + 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    xs = [0, 1, 2, 7, 14, 15, 12345, 2 ** 32, 2 ** 63, M]
    with open(os.path.join(here, "splitmix64_kat.json"), "w") as f:
        json.dump({"pairs": [[str(x), f"{splitmix64(x):#018x}"] for x in xs]}, f, indent=1)
