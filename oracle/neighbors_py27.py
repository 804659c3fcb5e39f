"""Pure-Python restatement of tools.GenerateNeighbors (piqmc/tools.pyx:28-96).

TEST INFRASTRUCTURE ONLY (the checker for the product's C++ builder).

The reference iterates ``J.todok().iterkeys()`` (piqmc/tools.pyx:79-84) under
CPython 2.7, so the order in which neighbours land in each row is CPython 2.7
dict iteration order: 64-bit tuple hash (Objects/tupleobject.c ``tuplehash``),
open addressing with the ``i = 5*i + perturb + 1; perturb >>= 5`` probe
(Objects/dictobject.c ``lookdict``), resize to the smallest power of two above
4*used (2*used past 50000) whenever fill*3 >= 2*(mask+1), rehashing the old
table in slot order.  Those published CPython 2.7 algorithms are restated here;
the result is pinned by the golden array of testing/test_core.py:53-78.
"""
import numpy as np

_M64 = (1 << 64) - 1


def _signed(x):
    x &= _M64
    return x - (1 << 64) if x >> 63 else x


def _int_hash(v):
    h = int(v)
    return -2 if h == -1 else h


def tuple_hash(t):
    x = 0x345678
    mult = 1000003
    n = len(t)
    for item in t:
        n -= 1
        x = _signed((x ^ _int_hash(item)) * mult)
        mult = _signed(mult + 82520 + n + n)
    x = _signed(x + 97531)
    return -2 if x == -1 else x


def _probe(table, mask, h, key):
    i = h & _M64
    perturb = h & _M64
    slot = i & mask
    while table[slot] is not None and table[slot][1] != key:
        i = (5 * i + perturb + 1) & _M64
        slot = i & mask
        perturb >>= 5
    return slot


def py27_order(keys):
    """Iteration order of a CPython 2.7 dict after inserting ``keys`` in order."""
    mask, table, used, fill = 7, [None] * 8, 0, 0
    for key in keys:
        key = (int(key[0]), int(key[1]))
        h = tuple_hash(key)
        slot = _probe(table, mask, h, key)
        if table[slot] is not None:
            continue                       # existing key: order unchanged
        table[slot] = (h, key)
        used += 1
        fill += 1
        if fill * 3 >= (mask + 1) * 2:
            minused = (2 if used > 50000 else 4) * used
            size = 8
            while size <= minused:
                size <<= 1
            old, mask, table = table, size - 1, [None] * size
            for ent in old:
                if ent is not None:
                    table[_probe(table, mask, ent[0], ent[1])] = ent
            fill = used
    return [ent[1] for ent in table if ent is not None]


def generate_neighbors(nspins, keys, values, maxnb):
    """Restates piqmc/tools.pyx:73-92 over (key, value) pairs in DOK
    insertion order.  Raises IndexError on a row overflow, as the
    bounds-checked Cython buffer access at tools.pyx:86-91 does."""
    val = {}
    for k, v in zip(keys, values):
        val[(int(k[0]), int(k[1]))] = float(v)
    order = py27_order(list(keys))
    nbs = np.zeros((nspins, maxnb, 2))
    for ispin in range(nspins):
        ipair = 0
        for pair in order:
            if pair[0] == ispin or pair[1] == ispin:
                if ipair >= maxnb:
                    raise IndexError("Out of bounds on buffer access (axis 1)")
                other = pair[1] if pair[0] == ispin else pair[0]
                nbs[ispin, ipair, 0] = other
                nbs[ispin, ipair, 1] = val[pair]
                ipair += 1
    return nbs
