"""Inputs of the reference's propagation harness (src/cuda_fitness.rs:537-539):
`Xoshiro256StarStar::seed_from_u64(0)` (rand_xoshiro 0.6) and rand 0.8's
`Standard` f32 samples, restated so the per-ray test traces the same 1024
(wavelength cdf, ray cdf) pairs as the reference's own test.

- seed_from_u64 fills the 256-bit state with four SplitMix64 outputs (state =
  the seed), little-endian words in order;
- xoshiro256** (Blackman and Vigna): result = rotl(s1 * 5, 7) * 9, then the
  xorshift state update with rotl(s3, 45);
- next_u32 takes the upper half of next_u64 (the low bits of the xoshiro
  family have linear dependencies);
- Standard f32 = (next_u32 >> 8) * 2^-24, in [0, 1).

Test infrastructure: pinned against the published xoshiro256** and SplitMix64
outputs in tests/test_host.py."""
import numpy as np

M64 = (1 << 64) - 1


def splitmix64(state: int):
    """(next state, output) of one SplitMix64 step."""
    state = (state + 0x9E3779B97F4A7C15) & M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return state, z ^ (z >> 31)


def _rotl(x: int, k: int) -> int:
    return ((x << k) | (x >> (64 - k))) & M64


class Xoshiro256StarStar:
    def __init__(self, s):
        self.s = [int(v) & M64 for v in s]

    @classmethod
    def seed_from_u64(cls, seed: int):
        st, words = seed & M64, []
        for _ in range(4):
            st, z = splitmix64(st)
            words.append(z)
        return cls(words)

    def next_u64(self) -> int:
        s = self.s
        result = (_rotl((s[1] * 5) & M64, 7) * 9) & M64
        t = (s[1] << 17) & M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = _rotl(s[3], 45)
        return result

    def next_u32(self) -> int:
        return self.next_u64() >> 32

    def standard_f32(self, n: int) -> np.ndarray:
        return np.array([(self.next_u32() >> 8) * (1.0 / (1 << 24)) for _ in range(n)], np.float32)


def harness_pairs(n: int = 1024):
    """The reference harness's wavelength cdf and ray cdf arrays (n each, drawn
    in that order from one generator)."""
    rng = Xoshiro256StarStar.seed_from_u64(0)
    w = rng.standard_f32(n)
    r = rng.standard_f32(n)
    return w, r
