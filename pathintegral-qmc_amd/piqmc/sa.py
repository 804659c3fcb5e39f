"""Drop-in for piqmc.sa (piqmc/sa.pyx): classical annealing and energies.

``Anneal`` runs the replay kernel with one slice (per-spin reset, ``>=``
acceptance, float temperature per step, piqmc/sa.pyx:50-120) and reproduces
the Cython code bit-for-bit, drawing the same permutations from ``rng`` and
the same libc ``rand()`` stream.  ``Anneal_dense`` (piqmc/sa.pyx:126-187) is
the same for a dense coupling matrix.  ``ClassicalIsingEnergy`` evaluates
piqmc/sa.pyx:41-44 on the GPU.
"""
import ctypes as C

import numpy as np

from . import _crand, _lib
from ._args import c_int, elem_strides, f64_view

last_near_ties = 0
last_draws = 0


def _coo(J):
    if hasattr(J, "tocoo"):
        coo = J.tocoo()
        return (np.asarray(coo.row, np.int64), np.asarray(coo.col, np.int64),
                np.asarray(coo.data, np.float64), coo.shape[0])
    J = np.asarray(J, np.float64)
    r, c = np.nonzero(J)
    return r.astype(np.int64), c.astype(np.int64), J[r, c].astype(np.float64), J.shape[0]


def ClassicalIsingEnergy(spins, J):
    """ClassicalIsingEnergy(spins, J)

    Energy -s.(J_off s) - sum(diag(J) s) of configuration ``spins`` under the
    coupling matrix ``J`` (scipy sparse or dense; duplicates add), as
    piqmc/sa.pyx:25-44.  A 2-D ``spins`` [B, N] returns B energies.
    """
    rows, cols, vals, n = _coo(J)
    s = np.ascontiguousarray(np.asarray(spins, dtype=np.float64))
    single = s.ndim == 1
    s2 = s[None, :] if single else s
    if s2.shape[1] != n:
        raise ValueError("spins length %d does not match J (%d)" % (s2.shape[1], n))
    E = np.zeros(s2.shape[0])
    _lib.call("piqmc_classical_energy", n, len(vals), rows.ctypes.data, cols.ctypes.data,
              vals.ctypes.data, s2.shape[0], s2.ctypes.data, E.ctypes.data)
    return np.float64(E[0]) if single else E


def Anneal(sched, mcsteps, svec, nbs, rng):
    """Anneal(sched, mcsteps, svec, nbs, rng)

    Thermal annealing of ``svec`` (+1/-1 float64) in place following the
    temperature schedule ``sched``, piqmc/sa.pyx:50-120.  Returns None.
    """
    return _anneal(sched, mcsteps, svec, nbs, rng, dense=False)


def Anneal_dense(sched, mcsteps, svec, J, rng):
    """Anneal_dense(sched, mcsteps, svec, J, rng)

    Anneal with a dense coupling matrix ``J`` (upper triangle and diagonal
    read, in double precision), piqmc/sa.pyx:126-187.  Unlike ``Anneal`` it
    accepts a move only when ediff > 0 before the Metropolis draw
    (piqmc/sa.pyx:181).  Returns None.
    """
    return _anneal(sched, mcsteps, svec, J, rng, dense=True)


def _anneal(sched, mcsteps, svec, cpl, rng, dense):
    global last_near_ties, last_draws
    sched = f64_view(sched, 1, "sched")
    svec = f64_view(svec, 1, "svec", writable=True)
    cpl = f64_view(cpl, 2 if dense else 3, "J" if dense else "nbs")
    mcsteps = c_int(mcsteps)
    nspins = svec.shape[0]
    S = sched.shape[0]
    nsweeps = S * max(mcsteps, 0)
    perm = rng.permutation(range(nspins))
    if S == 0 or mcsteps <= 0:
        return None
    if nspins == 0:
        for _ in range(nsweeps):
            perm = rng.permutation(perm)
        return None
    if dense:
        if cpl.shape[0] < nspins or cpl.shape[1] < nspins:
            raise ValueError("J must be at least (nspins, nspins) = (%d, %d), got %s"
                             % (nspins, nspins, cpl.shape))
    elif cpl.shape[0] != nspins:
        raise ValueError("nbs must have one row per spin")
    perms = np.empty((nsweeps, nspins), np.int64)
    for g in range(nsweeps):
        perms[g] = perm
        perm = rng.permutation(perm)
    sched_c = np.ascontiguousarray(sched)
    st = _crand.get_state()
    draws, ties = C.c_uint64(0), C.c_uint32(0)
    if dense:
        js0, js1 = elem_strides(cpl)
        _lib.call("piqmc_anneal_dense", sched_c.ctypes.data, S, mcsteps, svec.ctypes.data,
                  svec.strides[0] // 8, nspins, cpl.ctypes.data, js0, js1,
                  perms.ctypes.data, st.ctypes.data, C.byref(draws), C.byref(ties))
    else:
        nbs_c = np.ascontiguousarray(cpl)
        _lib.call("piqmc_anneal", sched_c.ctypes.data, S, mcsteps, svec.ctypes.data,
                  svec.strides[0] // 8, nspins, nbs_c.ctypes.data, nbs_c.shape[1],
                  perms.ctypes.data, st.ctypes.data, C.byref(draws), C.byref(ties))
    _crand.set_state(st)
    last_draws, last_near_ties = draws.value, ties.value
    return None
