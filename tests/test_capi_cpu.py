"""CPU-side checks of the product: the C ABI loads and exports every declared
symbol, the host services (GenerateNeighbors, the reference's random streams,
the libc state bridge) match their oracles, and argument errors raise the
reference's exception types before any device work."""
import ctypes
import os
import re

import numpy as np
import pytest
import scipy.sparse as sps

import oracle as O
from conftest import ROOT, boixo_J

libc = ctypes.CDLL("libc.so.6")


def test_library_exports_every_declared_symbol():
    from piqmc import _lib
    hdr = open(os.path.join(ROOT, "include", "piqmc_amd.h")).read()
    declared = set(re.findall(r"\b(piqmc_[a-z0-9_]+)\s*\(", hdr))
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(_lib.lib, name), name
    assert declared == set(_lib.EXPORTED)


# ------------------------------------------------------------ tools (testing/test_core.py)
def test_spinbitconversion():
    from piqmc import tools
    a = np.array([0., 1., 0., 1., 0., 1., 0., 1., 0., 1.])
    b = tools.bits2spins(a)
    c = tools.spins2bits(b)
    assert np.all(a == [0, 1, 0, 1, 0, 1, 0, 1, 0, 1])
    assert np.all(np.array(b) == [1, -1, 1, -1, 1, -1, 1, -1, 1, -1])
    assert np.all(np.array(c) == a)


def test_generateneighbors_golden():
    # testing/test_core.py:30-81 (CPython 2.7 dict order of the DOK keys)
    from piqmc import tools
    J = sps.dok_matrix((5, 5), dtype=np.float64)
    for k in [(0, 1), (1, 2), (2, 3), (3, 4), (0, 3), (1, 4), (0, 4), (0, 0), (1, 1), (2, 2),
              (3, 3), (4, 4)]:
        J[k] = 1
    true_nb = np.array([[[1., 1.], [0., 1.], [4., 1.], [3., 1.]],
                        [[0., 1.], [2., 1.], [4., 1.], [1., 1.]],
                        [[1., 1.], [2., 1.], [3., 1.], [0., 0.]],
                        [[3., 1.], [2., 1.], [0., 1.], [4., 1.]],
                        [[4., 1.], [1., 1.], [0., 1.], [3., 1.]]])
    assert np.all(true_nb == tools.GenerateNeighbors(nspins=5, J=J, maxnb=4, savepath=None))


@pytest.mark.parametrize("seed", range(6))
def test_generateneighbors_matches_py27_oracle(seed):
    from neighbors_py27 import generate_neighbors
    from piqmc import tools
    rng = np.random.RandomState(seed)
    N = int(rng.randint(5, 300))
    J = sps.dok_matrix((N, N))
    keys = []
    for _ in range(int(N * rng.uniform(1, 3))):
        i, j = sorted(rng.randint(0, N, 2))
        if (i, j) not in J.keys():
            keys.append((i, j))
        J[i, j] = rng.uniform(-2, 2)
    deg = np.zeros(N, int)
    for i, j in J.keys():
        deg[i] += 1
        deg[j] += (i != j)
    maxnb = int(deg.max()) + int(rng.randint(0, 3))
    ref = generate_neighbors(N, list(J.keys()), [J[k] for k in J.keys()], maxnb)
    assert np.array_equal(tools.GenerateNeighbors(N, J, maxnb), ref)


def test_generateneighbors_overflow_and_save(tmp_path):
    from piqmc import tools
    J = boixo_J()
    with pytest.raises(IndexError):
        tools.GenerateNeighbors(8, J, 2)
    p = str(tmp_path / "nbs.npy")
    nb = tools.GenerateNeighbors(8, J, 4, savepath=p)
    assert np.array_equal(np.load(p), nb)


def test_lattice_generator_upper_triangular():
    # the convention of testing/test_core.py:24-28
    from piqmc import instances
    N, r, c, v = instances.lattice_pm_j(4, 2, np.random.RandomState(123))
    J = sps.coo_matrix((v, (r, c)), shape=(N, N))
    assert (J - sps.triu(J)).nnz == 0
    assert set(np.unique(v)) <= {-1.0, 1.0} and len(v) == 2 * N


# ------------------------------------------------------------ streams
def test_rand_seed_matches_libc():
    from piqmc import device
    for seed in [1, 5, 2**31 + 7]:
        assert np.array_equal(device.rand_seed(seed), O.Type3(seed).st)


def test_perm_streams_match_numpy():
    from piqmc import device
    seeds = [11, 12, 13, 2**32 - 1]
    host = device.perm_streams(seeds, 50, 6)
    for r, s in enumerate(seeds):
        rng = np.random.RandomState(s)
        p = [rng.permutation(range(50))]
        for _ in range(5):
            p.append(rng.permutation(p[-1]))
        assert np.array_equal(host[:, :, r], np.array(p))


def test_libc_state_bridge_roundtrip():
    from piqmc import _crand
    libc.srand(777)
    t = O.Type3(777)
    assert np.array_equal(_crand.get_state(), t.st)
    for _ in range(123):
        t.next()
    _crand.set_state(t.st)
    assert [libc.rand() for _ in range(5)] == [t.next() for _ in range(5)]
    _crand.srand(31)
    assert libc.rand() == O.Type3(31).next()


# ------------------------------------------------------------ argument errors
def test_dropin_argument_errors():
    from piqmc import qmc, sa
    nbs = np.zeros((4, 2, 2))
    confs = np.ones((4, 3))
    sched = np.linspace(1, 0.1, 3)
    rng = np.random.RandomState(0)
    with pytest.raises(ValueError):          # Buffer dtype mismatch
        qmc.QuantumAnneal(sched.astype(np.float32), 1, 3, 0.1, 4, confs, nbs, rng)
    with pytest.raises(ValueError):          # wrong rank
        qmc.QuantumAnneal(sched, 1, 3, 0.1, 4, confs[:, 0], nbs, rng)
    with pytest.raises(TypeError):
        qmc.QuantumAnneal([1.0, 0.5], 1, 3, 0.1, 4, confs, nbs, rng)
    with pytest.raises(ZeroDivisionError):   # slices*temp == 0, qmc.c:2055-2075
        qmc.QuantumAnneal(sched, 1, 3, 0.0, 4, confs, nbs, rng)
    with pytest.raises(ValueError):          # slice 1 does not exist
        qmc.QuantumAnneal(sched, 1, 1, 0.1, 4, np.ones((4, 1)), nbs, rng)
    with pytest.raises(ValueError):          # non +-1 spins
        qmc.QuantumAnneal(sched, 1, 3, 0.1, 4, confs * 0.5, nbs, rng)
    ro = confs.copy()
    ro.setflags(write=False)
    with pytest.raises(ValueError):
        qmc.QuantumAnneal(sched, 1, 3, 0.1, 4, ro, nbs, rng)
    with pytest.raises(ValueError):
        sa.Anneal(sched, 1, np.ones(4, np.int64), nbs, rng)


def test_problem_rejects_bad_tables():
    from piqmc import device
    bad = np.zeros((3, 1, 2))
    bad[1, 0] = [5, 1.0]                    # neighbour index beyond nspins
    with pytest.raises(ValueError):
        device.Problem(bad)
