"""Writes the known-answer vectors under tests/golden/.

The reference ships no fixtures for the PatchMatch path, so the vectors
here come from the published definitions, computed independently of the
oracle with Python integers:

* mwc64x_kat.json -- MWC64X (D. B. Thomas 2011, vendored as
  cvt/cl/kernel/MWC64X.cl): with A = 4294883355, M = A*2^32 - 1 and the
  seed constant BASE = 4077358422479273989, SeedStreams(0, 0) + Skip(k)
  puts the generator at v = BASE * A^k mod M, state x = v // A,
  c = v % A (MWC64X.cl:78-107), and RNG_float returns
  float32(x ^ c) * 2^-32 (RNG.cl:30-35).  Besides the jump formula the
  file holds the first 64 states reached by literally iterating
  MWC64X_Step (MWC64X.cl:109-120) from the seed, which pins the jump
  against the step.

Run: python tests/golden/make_golden.py
"""
import json
import os

import numpy as np

A = 4294883355
M = A * 2 ** 32 - 1
BASE = 4077358422479273989
HERE = os.path.dirname(os.path.abspath(__file__))


def state_at(k):
    v = BASE * pow(A, k, M) % M
    return v // A, v % A


def draw(x, c):
    return float(np.float32(np.float32(x ^ c) * np.float32(2.0 ** -32)))


def step(x, c):
    t = A * x + c
    return t % 2 ** 32, t // 2 ** 32


def main():
    ks = list(range(0, 16)) + [3 * 1000 + 2, 14 * 5 + 13, 14 * (1280 * 720)
                               + 13, 2 ** 31 - 1, 2 ** 32 + 5, 2 ** 40 + 7,
                               14 * (1920 * 1080 + 7)]
    jump = []
    for k in ks:
        x, c = state_at(k)
        jump.append({"k": k, "x": x, "c": c, "draw": draw(x, c)})
    x, c = state_at(0)
    stepped = []
    for i in range(64):
        stepped.append({"k": i, "x": x, "c": c, "draw": draw(x, c)})
        x, c = step(x, c)
    out = {"A": A, "M": M, "BASE": BASE, "jump": jump, "stepped": stepped}
    with open(os.path.join(HERE, "mwc64x_kat.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
