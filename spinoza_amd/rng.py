"""Seeded uniform draws for measurement, noise and sampling.

The reference draws from an unseeded ``thread_rng`` (measurement.rs:35-36,
core.rs:88-109, gates.rs:1366).  Here every draw comes from SplitMix64
(Steele, Lea and Flood, 2014), u = (x >> 11) * 2^-53, so a run is
reproducible from its seed and the same uniforms can be handed to the oracle.
"""
import os

import numpy as np

_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed=0):
        self.state = seed & _MASK

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def uniforms(self, count):
        """`count` uniforms in [0, 1) (vectorised, same stream as uniform())."""
        if count <= 0:
            return np.zeros(0)
        start = np.uint64(self.state)
        inc = np.uint64(0x9E3779B97F4A7C15)
        with np.errstate(over="ignore"):
            z = start + inc * np.arange(1, count + 1, dtype=np.uint64)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        self.state = (self.state + 0x9E3779B97F4A7C15 * count) & _MASK
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def uniforms(seed, count):
    """`count` uniforms of the SplitMix64 stream that starts at `seed`."""
    return SplitMix64(seed).uniforms(count)


def gaussian_state(seed, n):
    """(re, im) of a normalised 2^n-amplitude state with N(0, 1) parts
    (Box-Muller on consecutive SplitMix64 pairs: u1 = 1 - u_{2i}, u2 = u_{2i+1},
    re = r cos(2 pi u2), im = r sin(2 pi u2), r = sqrt(-2 ln u1))."""
    u = SplitMix64(seed).uniforms(2 << n)
    u1 = 1.0 - u[0::2]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    re = r * np.cos(2.0 * np.pi * u2)
    im = r * np.sin(2.0 * np.pi * u2)
    scale = 1.0 / np.sqrt(np.sum(re * re) + np.sum(im * im))
    return re * scale, im * scale


_default = SplitMix64(int(os.environ.get("SPINOZA_SEED", "0x5EED"), 0))


def default_rng():
    return _default


def seed(value):
    """Reseed the process-wide generator."""
    _default.state = int(value) & _MASK
