"""Writes tests/golden/generate.json: raw bits of the first values of the
seeded generator for a few seeds and all distributions, computed with
Python integers and exact rationals (independent of the C oracle and the
GPU kernel).  Run: python tests/golden/make_generate.py"""
import json
import os
import struct
from fractions import Fraction

M64 = (1 << 64) - 1


def splitmix(seed, n):
    s, out = seed, []
    for _ in range(n):
        s = (s + 0x9E3779B97F4A7C15) & M64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        out.append(z ^ (z >> 31))
    return out


def bits(x):
    return "%016x" % struct.unpack("<Q", struct.pack("<d", float(x)))[0]


def values(seed, dist, n):
    zs = splitmix(seed, n)
    ze = splitmix(seed ^ 0xD1B54A32D192ED03, n)
    out = []
    for i in range(n):
        u = Fraction(zs[i] >> 11, 1 << 53)
        if dist == 2:
            v = u
        else:
            v = 2 * u - 1
            if dist == 1:
                v = v * Fraction(2) ** ((ze[i] >> 32) % 41 - 20)
        out.append(bits(v))  # exact: every value is a dyadic rational with <= 53 significant bits
    return out


if __name__ == "__main__":
    cases = [{"seed": s, "dist": d, "bits": values(s, d, 16)} for s in (1, 2, 3, 42) for d in (0, 1, 2)]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "generate.json")
    with open(path, "w") as f:
        json.dump({"generator": "SplitMix64 counter-based, see include/dgamd.h dg_generate", "cases": cases}, f,
                  indent=1)
    print("wrote", path)
