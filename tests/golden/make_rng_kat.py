"""Writes tests/golden/rng_kat.json: known answers for LiSA's RNG.

Computed from the published algorithms, independently of the oracle:
* TEA (Wheeler & Needham, 1994) with a 128-bit key and N rounds; the
  generic function is first checked against the published TEA vector
  (key 0, block 0, 32 rounds -> 0x41ea3a0a 0x94baa940).  LiSA's tea<16>
  (src/cuda/random.h:31-46) is TEA with 16 rounds and the key
  (0xa341316c, 0xc8013ea4, 0xad90777d, 0x7e95761e), returning v0.
* The Numerical Recipes LCG x' = 1664525 x + 1013904223 mod 2^32
  (random.h:49-55), whose first states from 0 are published in
  Numerical Recipes (1013904223, 1196435762, 3519870697, 2868466484).
"""
import json
import os

M = 0xFFFFFFFF


def tea(v0, v1, key, rounds):
    s = 0
    for _ in range(rounds):
        s = (s + 0x9E3779B9) & M
        v0 = (v0 + ((((v1 << 4) & M) + key[0]) & M ^ (v1 + s) & M ^ ((v1 >> 5) + key[1]) & M)) & M
        v1 = (v1 + ((((v0 << 4) & M) + key[2]) & M ^ (v0 + s) & M ^ ((v0 >> 5) + key[3]) & M)) & M
    return v0, v1


LISA_KEY = (0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E)


def lcg_states(seed, n):
    out = []
    for _ in range(n):
        seed = (1664525 * seed + 1013904223) & M
        out.append(seed)
    return out


def main():
    assert tea(0, 0, (0, 0, 0, 0), 32) == (0x41EA3A0A, 0x94BAA940)
    assert lcg_states(0, 4) == [1013904223, 1196435762, 3519870697, 2868466484]
    pairs = [(0, 0), (1, 0), (0, 1), (255, 0), (65535, 3), (1048575, 0), (123456, 7),
             (2073599, 0), (0xFFFFFFFF, 0xFFFFFFFF)]
    tea16 = [[a, b, tea(a, b, LISA_KEY, 16)[0]] for a, b in pairs]
    seeds = [0, 1, 0xDEADBEEF, tea16[3][2]]
    lcg = [[s, lcg_states(s, 8)] for s in seeds]
    rnd = [[s, [(x & 0xFFFFFF) / float(1 << 24) for x in lcg_states(s, 8)]] for s in seeds]
    out = {"tea16": tea16, "lcg_states": lcg, "rnd": rnd}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng_kat.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
