"""Writes the small golden fixtures used by the CPU tests.

* splitmix64.json -- first outputs of SplitMix64 for a few seeds, computed here
  with Python integers (an implementation independent of the oracle's C code),
  plus the published test value for seed 1234567.
* examples/*.toml -- the reference's example job files (data, copied verbatim
  from examples/ of the reference) used as parser fixtures.
Run from the repo root: python3 tests/golden/make_golden.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
M = (1 << 64) - 1
G = 0x9E3779B97F4A7C15


def splitmix64(seed, n):
    s = seed & M
    out = []
    for _ in range(n):
        s = (s + G) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        out.append(z ^ (z >> 31))
    return out


if __name__ == "__main__":
    data = {str(s): [str(v) for v in splitmix64(s, 8)]
            for s in (0, 1, 11, 12345, 1234567, 20260710, M)}
    with open(os.path.join(HERE, "splitmix64.json"), "w") as f:
        json.dump(data, f, indent=1)
    print("wrote splitmix64.json")
