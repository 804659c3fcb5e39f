"""Ising instances: the +-J lattices, SK glasses and text format the sweeps run on.

The reference ships instances as text files of 1-based ``i j J`` lines
(examples/ising_instances/gen_instances.py:71-78, read by
examples/spinglass32.py:81-84) and generates only U[-2,2] lattices
(piqmc/tools.pyx:98-130).  The workloads here need +-J lattices in 2-D and 3-D
with optional Gaussian fields and dense Sherrington-Kirkpatrick couplings;
all generators return an upper-triangular COO list (diagonal = local fields),
the convention of testing/test_core.py:24-28.
"""
import numpy as np
import scipy.sparse as sps

from . import tools


def lattice_pm_j(L, dim=2, rng=None, field_sigma=0.0):
    """Periodic L^dim hypercubic lattice with J_ij = +-1 equiprobable on the
    +x/+y(/+z) bonds, optional fields h_i ~ N(0, field_sigma^2) (float32
    values).  Returns (N, rows, cols, vals)."""
    rng = np.random.RandomState(rng) if not isinstance(rng, np.random.RandomState) else rng
    N = L ** dim
    idx = np.arange(N).reshape((L,) * dim)
    rows, cols = [], []
    for ax in range(dim):
        nb = np.roll(idx, -1, axis=ax)
        rows.append(idx.ravel())
        cols.append(nb.ravel())
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    vals = (2 * rng.randint(2, size=rows.size) - 1).astype(np.float64)
    lo, hi = np.minimum(rows, cols), np.maximum(rows, cols)
    if field_sigma:
        h = rng.normal(0.0, field_sigma, N).astype(np.float32).astype(np.float64)
        lo = np.concatenate([lo, np.arange(N)])
        hi = np.concatenate([hi, np.arange(N)])
        vals = np.concatenate([vals, h])
    return N, lo.astype(np.int64), hi.astype(np.int64), vals


def sk(N, rng=None):
    """Dense Sherrington-Kirkpatrick couplings J_ij ~ N(0, 1/N) (float32
    values), i < j.  Returns (N, rows, cols, vals)."""
    rng = np.random.RandomState(rng) if not isinstance(rng, np.random.RandomState) else rng
    r, c = np.triu_indices(N, 1)
    vals = rng.normal(0.0, 1.0 / np.sqrt(N), r.size).astype(np.float32).astype(np.float64)
    return N, r.astype(np.int64), c.astype(np.int64), vals


def to_matrix(N, rows, cols, vals):
    return sps.coo_matrix((vals, (rows, cols)), shape=(N, N)).todok()


def neighbors(N, rows, cols, vals, maxnb=None, order="sorted"):
    """Neighbour table for a COO instance (tools.GenerateNeighbors)."""
    if maxnb is None:
        deg = np.bincount(rows, minlength=N) + np.bincount(cols[cols != rows], minlength=N)
        maxnb = int(deg.max()) if N else 1
    J = sps.coo_matrix((vals, (rows, cols)), shape=(N, N))
    return tools.GenerateNeighbors(N, J, maxnb, order=order)


def read_instance(path):
    """1-based ``i j J`` text file -> (N, rows, cols, vals) (0-based)."""
    a = np.loadtxt(path, ndmin=2)
    rows = a[:, 0].astype(np.int64) - 1
    cols = a[:, 1].astype(np.int64) - 1
    N = int(max(rows.max(), cols.max())) + 1
    return N, rows, cols, a[:, 2].astype(np.float64)


def write_instance(path, rows, cols, vals):
    with open(path, "w") as f:
        for i, j, v in zip(rows, cols, vals):
            f.write("%d %d %.12f \n" % (i + 1, j + 1, v))


def random_spins(N, rng):
    """+-1 initial spins as the examples draw them (2*rng.randint(2)-1 per
    spin, examples/spinglass32.py:94-95)."""
    return np.array([2 * rng.randint(2) - 1 for _ in range(N)], dtype=np.float64)


def linear_schedule(start, stop, steps):
    return np.linspace(start, stop, steps)
