"""Counter-based SplitMix64 streams used for every seeded synthetic input.

Value ``j`` of stream ``seed`` is ``mix(seed + (j + 1) * GOLDEN)`` where ``mix``
is Steele/Lea/Flood's SplitMix64 finaliser (OOPSLA 2014).  Being counter based,
any element can be generated independently, which keeps the numpy generator
vectorised and lets C/HIP code reproduce the same numbers.
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, count, start=0):
    """Return ``count`` uint64 values of stream ``seed`` starting at index ``start``."""
    with np.errstate(over="ignore"):
        j = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + j * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, count, start=0):
    """Uniform doubles in [0, 1) with 53 random bits each."""
    return (splitmix64(seed, count, start) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed, count, start=0):
    """Standard normal deviates by the Box-Muller transform of two uniform streams."""
    u1 = uniform(seed, count, start)
    u2 = uniform(seed ^ 0x5DEECE66D, count, start)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)
