"""Writes the golden fixtures under tests/golden/.

rng_pairs.npz: start-vector uniforms of vector 7 for plane waves 1..64 from
the SplitMix64 finaliser evaluated with Python integers (an implementation
independent of the vectorised NumPy one in oracle/davidson.py and of the HIP
one in the library):

    mix64(x) = z ^ (z >> 31), z = ((y ^ (y >> 27)) * 0x94D049BB133111EB),
    y = ((w ^ (w >> 30)) * 0xBF58476D1CE4E5B9), w = x + 0x9E3779B97F4A7C15  (mod 2^64)
    u_k = (mix64(SEED ^ (ivec << 32) ^ (2*ig + k)) >> 11) / 2^53
"""
from pathlib import Path

import numpy as np

M64 = (1 << 64) - 1
SEED = 0x5EED0F1E57


def mix64(x: int) -> int:
    w = (x + 0x9E3779B97F4A7C15) & M64
    y = ((w ^ (w >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((y ^ (y >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def uniforms(ivec: int, ig: int):
    base = SEED ^ (ivec << 32)
    return tuple((mix64(base ^ (2 * ig + k)) >> 11) / float(1 << 53) for k in (0, 1))


def main():
    out = Path(__file__).parent
    ig = np.arange(1, 65)
    u = np.array([uniforms(7, int(g)) for g in ig])
    np.savez(out / "rng_pairs.npz", ig=ig, u0=u[:, 0], u1=u[:, 1])


if __name__ == "__main__":
    main()
