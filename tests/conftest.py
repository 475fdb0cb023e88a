import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def uniform_box(n, lengths, seed_c=1001, seed_q=1002):
    """Charges uniform in a box with the SplitMix64 stream, q ~ U[-1,1]
    shifted to zero net charge (the input convention of the benchmarks)."""
    from oracle import oracle as O
    L = np.broadcast_to(np.asarray(lengths, dtype=np.float64), (3,))
    coord = O.splitmix_uniform(seed_c, 3 * n).reshape(n, 3) * L
    q = O.splitmix_uniform(seed_q, n) * 2.0 - 1.0
    q -= q.mean()
    return coord, q


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O
