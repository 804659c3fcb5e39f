import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "pathintegral-qmc_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def boixo_J():
    """The 8-spin Boixo graph of testing/test_boixo.py:34-54 (1-based list)."""
    import scipy.sparse as sps
    kvp = [[1, 2, 1.0], [1, 4, 1.0], [1, 5, 1.0], [2, 3, 1.0], [2, 6, 1.0], [3, 4, 1.0],
           [3, 7, 1.0], [4, 8, 1.0], [1, 1, 1.0], [2, 2, 1.0], [3, 3, 1.0], [4, 4, 1.0],
           [5, 5, -1.0], [6, 6, -1.0], [7, 7, -1.0], [8, 8, -1.0]]
    J = sps.dok_matrix((8, 8))
    for i, j, v in kvp:
        J[i - 1, j - 1] = v
    return J


def pm_j_problem(L, dim=2, seed=0, field_sigma=0.0):
    """(N, nbs, rows, cols, vals) for a periodic +-J lattice."""
    from piqmc import instances
    N, r, c, v = instances.lattice_pm_j(L, dim, np.random.RandomState(seed), field_sigma)
    nbs = instances.neighbors(N, r, c, v)
    return N, nbs, r, c, v


def init_confs(N, P, rng):
    """Random +-1 spins tiled over P slices, F-ordered like
    np.tile(spinVector, (P, 1)).T (examples/spinglass32.py:94-96)."""
    sv = np.array([2 * rng.randint(2) - 1 for _ in range(N)], dtype=np.float64)
    return np.tile(sv, (P, 1)).T
