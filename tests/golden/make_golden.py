"""Writes tests/golden/*.json: known-answer vectors for the oracle.

Every value here is computed by an independent restatement in Python
(arbitrary-precision integers and float64), not by the C oracle:
  rng_streams.json   random.wgsl:11-22 (u32 hash chain, f32 conversion)
  xoshiro.json       Zig 0.11 std.rand.SplitMix64 / Xoshiro256 (xoshiro256++)
                     and Random.uintLessThan, for the BVH split axis
The reference itself ships no test vectors (no tests at all), so these pin
the restatement against the published algorithms.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
M32 = 0xFFFFFFFF
M64 = (1 << 64) - 1


def rng_u32(s):
    x = (s + 747796405 + 2891336453) & M32
    w = (((x >> ((x >> 28) + 4)) ^ x) * 277803737) & M32
    return (w >> 22) ^ w


def rng_stream(state, n):
    out, s = [], state
    for _ in range(n):
        s = rng_u32(s)
        f = np.float32(s) * np.float32(1.0 / 4294967296.0)  # round-to-nearest u32 -> f32
        out.append((s, int(np.float32(f).view(np.uint32))))
    return out


def splitmix(state):
    state = (state + 0x9E3779B97F4A7C15) & M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return state, z ^ (z >> 31)


def rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & M64


def xoshiro_next(s):
    r = (rotl((s[0] + s[3]) & M64, 23) + s[0]) & M64
    t = (s[1] << 17) & M64
    s[2] ^= s[0]
    s[3] ^= s[1]
    s[1] ^= s[2]
    s[0] ^= s[3]
    s[2] ^= t
    s[3] = rotl(s[3], 45)
    return r


def less_than(s, bound):
    x = xoshiro_next(s)
    m = x * bound
    lo = m & M64
    if lo < bound:
        t = (-bound) % (1 << 64)
        if t >= bound:
            t -= bound
            if t >= bound:
                t %= bound
        while lo < t:
            x = xoshiro_next(s)
            m = x * bound
            lo = m & M64
    return m >> 64


def main():
    states = [0, 1, 2, 12345, 2 ** 31, 2 ** 32 - 1]
    rng = {str(s): rng_stream(s, 16) for s in states}
    # a state whose next draw converts to exactly 1.0f (u32 within 128 of 2^32)
    one = None
    for s in range(0, 1 << 26):
        if rng_u32(s) >= 0xFFFFFF80:
            one = s
            break
    json.dump({"streams": rng, "state_drawing_one": one}, open(os.path.join(HERE, "rng_streams.json"), "w"),
              indent=0)

    st, seq = 1244, []
    s = []
    for _ in range(4):
        st, z = splitmix(st)
        s.append(z)
    seeded = list(s)
    for _ in range(8):
        seq.append(xoshiro_next(s))
    s = list(seeded)
    axes = [less_than(s, 3) for _ in range(64)]
    sm0 = []
    st = 0
    for _ in range(4):
        st, z = splitmix(st)
        sm0.append(z)
    json.dump({"splitmix64_seed0": [str(v) for v in sm0],
               "xoshiro_1244_state": [str(v) for v in seeded],
               "xoshiro_1244_next": [str(v) for v in seq],
               "axes_1244": axes}, open(os.path.join(HERE, "xoshiro.json"), "w"), indent=0)


if __name__ == "__main__":
    main()
