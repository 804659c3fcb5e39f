"""Drop-in for piqmc.tools (piqmc/tools.pyx): the neighbour table builder.

``GenerateNeighbors`` produces the float64 [nspins, maxnb, 2] table the sweeps
consume.  By default the neighbours of each spin land in the order the
reference produces under CPython 2.7 (its ``J.todok().iterkeys()`` order,
piqmc/tools.pyx:79-92); ``order="insertion"`` and ``order="sorted"`` give the
DOK insertion order and row-major order.  The table is built in C++
(piqmc_generate_neighbors) in O(nnz) instead of O(nspins*nnz).
"""
import numpy as np

from . import _lib

_ORDERS = {"py27": 0, "insertion": 1, "sorted": 2}


def bits2spins(vec):
    """ Convert a bitvector @vec to a spinvector. (piqmc/tools.pyx:20-22) """
    return [-1 if k == 1 else 1 for k in vec]


def spins2bits(vec):
    """ Convert a spinvector @vec to a bitvector. (piqmc/tools.pyx:24-26) """
    return [0 if k == 1 else 1 for k in vec]


def _keys_in_insertion_order(J):
    """DOK keys in insertion order (a dict's order); other formats go through
    tocoo(), which is how J.todok() filled its dict in the reference era."""
    if hasattr(J, "format") and J.format == "dok":
        items = list(J.items())
        rows = np.array([k[0] for k, _ in items], np.int64)
        cols = np.array([k[1] for k, _ in items], np.int64)
        vals = np.array([v for _, v in items], np.float64)
        return rows, cols, vals
    if hasattr(J, "tocoo"):
        coo = J.tocoo()
        return (np.asarray(coo.row, np.int64), np.asarray(coo.col, np.int64),
                np.asarray(coo.data, np.float64))
    J = np.asarray(J, np.float64)
    r, c = np.nonzero(J)
    return r.astype(np.int64), c.astype(np.int64), J[r, c]


def GenerateNeighbors(nspins, J, maxnb, savepath=None, order="py27"):
    """GenerateNeighbors(nspins, J, maxnb, savepath=None)

    Neighbour table [nspins, maxnb, 2] of (neighbour index, coupling) rows; a
    diagonal entry (i, i) becomes a local-field row of spin i; unused rows
    stay [0, 0].  Raises IndexError when a spin has more than ``maxnb``
    entries (piqmc/tools.pyx:28-96).
    """
    nspins, maxnb = int(nspins), int(maxnb)
    rows, cols, vals = _keys_in_insertion_order(J)
    nbs = np.zeros((nspins, maxnb, 2), np.float64)
    _lib.call("piqmc_generate_neighbors", nspins, maxnb, len(vals), rows.ctypes.data,
              cols.ctypes.data, vals.ctypes.data, _ORDERS[order], nbs.ctypes.data)
    if savepath is not None:
        np.save(savepath, nbs)
    return nbs
