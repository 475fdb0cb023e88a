"""Seeded synthetic charge sets for benchmarks and tests.

SplitMix64 stream -> doubles in [0, 1) (the same generator as the oracle's
test inputs, so both sides see identical charges and coordinates)."""
import numpy as np

_M = (1 << 64) - 1


def splitmix_uniform(seed, n):
    i = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def uniform_box(n, lengths, seed_c=1001, seed_q=1002):
    """n charges uniform in a box, q ~ U[-1, 1] shifted to zero net charge."""
    L = np.broadcast_to(np.asarray(lengths, dtype=np.float64), (3,))
    coord = splitmix_uniform(seed_c, 3 * n).reshape(n, 3) * L
    q = splitmix_uniform(seed_q, n) * 2.0 - 1.0
    q -= q.mean()
    return coord, q


def metric_input(n=1000000, density=0.1):
    """The 1/r benchmark input: n charges uniform in a cube at the given
    density (side (n/density)^(1/3), 215.443 for n = 1M)."""
    L = (n / density) ** (1.0 / 3.0)
    return uniform_box(n, L)
