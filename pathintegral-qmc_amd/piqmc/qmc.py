"""Drop-in for piqmc.qmc (piqmc/qmc.pyx): path-integral quantum annealing.

``QuantumAnneal`` keeps the reference's signature and in-place semantics and
runs the replay kernel on the GPU, reproducing the Cython code bit-for-bit:
the same permutations are drawn from the caller's ``rng`` (1 + S*mcsteps
``rng.permutation`` calls, piqmc/qmc.pyx:90-91,136) and the same libc
``rand()`` stream is consumed and advanced (piqmc/qmc.pyx:132).

``QuantumAnneal_dense`` (piqmc/qmc.pyx:141-242) is the same drop-in for a
dense coupling matrix.

``QuantumAnneal_parallel`` keeps the signature of the reference's OpenMP
variant (piqmc/qmc.pyx:247-357) and runs the race-free graph-colored,
parity-alternated parallel sweep ("fixed" semantics: periodic Trotter
neighbours, per-spin reset) instead of the racy OpenMP loop.
"""
import ctypes as C

import numpy as np

from . import _crand, _lib
from ._args import c_float, c_int, elem_strides, f64_view

#: near-tie count of the last QuantumAnneal call (acceptance tests where the
#: device exp() and a correctly rounded exp() could decide differently).
last_near_ties = 0
#: libc rand() draws made by the last QuantumAnneal call.
last_draws = 0


def QuantumAnneal(sched, mcsteps, slices, temp, nspins, confs, nbs, rng):
    """QuantumAnneal(sched, mcsteps, slices, temp, nspins, confs, nbs, rng)

    Path-integral quantum annealing of one replica (P = slices Trotter
    slices) in place on ``confs`` [nspins, slices] (+1/-1 float64, any
    strides), piqmc/qmc.pyx:30-136.  Returns None.
    """
    return _quantum_anneal(sched, mcsteps, slices, temp, nspins, confs, nbs, rng, dense=False)


def QuantumAnneal_dense(sched, mcsteps, slices, temp, nspins, confs, J, rng):
    """QuantumAnneal_dense(sched, mcsteps, slices, temp, nspins, confs, J, rng)

    QuantumAnneal with a dense coupling matrix ``J`` (float64, off-diagonal
    couplings, diagonal fields; only the upper triangle is read) in place of
    the neighbour table, piqmc/qmc.pyx:141-242: the couplings enter the
    energy change in double precision and every spin's row is scanned in
    ascending order.  Returns None.
    """
    return _quantum_anneal(sched, mcsteps, slices, temp, nspins, confs, J, rng, dense=True)


def _quantum_anneal(sched, mcsteps, slices, temp, nspins, confs, cpl, rng, dense):
    global last_near_ties, last_draws
    sched = f64_view(sched, 1, "sched")
    confs = f64_view(confs, 2, "confs", writable=True)
    cpl = f64_view(cpl, 2 if dense else 3, "J" if dense else "nbs")
    mcsteps, slices, nspins = c_int(mcsteps), c_int(slices), c_int(nspins)
    temp = c_float(temp)
    S = sched.shape[0]
    nsweeps = S * max(mcsteps, 0)
    # Same order of rng use and exceptions as piqmc/qmc.pyx:90-136.
    perm = rng.permutation(range(max(nspins, 0)))
    if S == 0:
        return None
    if np.float32(slices) * np.float32(temp) == 0:
        raise ZeroDivisionError("float division")          # qmc.c:2055-2075
    if mcsteps <= 0:
        return None
    if slices < 0 or nspins < 1:                             # empty slice / spin loops
        for _ in range(nsweeps):
            perm = rng.permutation(perm)
        return None
    if confs.shape[0] != nspins or confs.shape[1] != slices:
        raise ValueError("confs must have shape (nspins, slices) = (%d, %d), got %s"
                         % (nspins, slices, confs.shape))
    if dense:
        if cpl.shape[0] < nspins or cpl.shape[1] < nspins:
            raise ValueError("J must be at least (nspins, nspins) = (%d, %d), got %s"
                             % (nspins, nspins, cpl.shape))
    elif cpl.shape[0] != nspins:
        raise ValueError("nbs must have nspins rows")
    if slices < 2:
        raise ValueError("QuantumAnneal needs slices >= 2 (the reference reads slice 1)")
    if np.float32(temp) == 0:
        raise ZeroDivisionError("float division")          # qmc.c:2400-2412
    perms = np.empty((nsweeps, nspins), np.int64)
    for g in range(nsweeps):
        perms[g] = perm
        perm = rng.permutation(perm)
    sched_c = np.ascontiguousarray(sched)
    st = _crand.get_state()
    draws, ties = C.c_uint64(0), C.c_uint32(0)
    cs0, cs1 = elem_strides(confs)
    if dense:
        js0, js1 = elem_strides(cpl)
        _lib.call("piqmc_quantum_anneal_dense", sched_c.ctypes.data, S, mcsteps, slices, temp,
                  nspins, confs.ctypes.data, cs0, cs1, cpl.ctypes.data, js0, js1,
                  perms.ctypes.data, st.ctypes.data, C.byref(draws), C.byref(ties))
    else:
        nbs_c = np.ascontiguousarray(cpl)
        _lib.call("piqmc_quantum_anneal", sched_c.ctypes.data, S, mcsteps, slices, temp, nspins,
                  confs.ctypes.data, cs0, cs1, nbs_c.ctypes.data, nbs_c.shape[1],
                  perms.ctypes.data, st.ctypes.data, C.byref(draws), C.byref(ties))
    _crand.set_state(st)
    last_draws, last_near_ties = draws.value, ties.value
    return None


_parallel_seeds = np.random.default_rng()


def QuantumAnneal_parallel(sched, mcsteps, slices, temp, nspins, confs, nbs, nthreads,
                           seed=None):
    """QuantumAnneal_parallel(sched, mcsteps, slices, temp, nspins, confs, nbs, nthreads)

    Parallel sweep on the GPU in place on ``confs`` [nspins, slices]
    (piqmc/qmc.pyx:247-357 signature; ``nthreads`` is accepted and ignored).
    ``seed`` keys the Philox stream; when None it comes from a generator
    private to this module, so numpy's global RandomState is left untouched.
    The sweep is the race-free graph-colored one: the reference's OpenMP
    variant is nondeterministic (racy reads of neighbouring spins), so no call
    can reproduce it.  Requires an even number of slices.
    """
    from . import device
    import torch

    sched = f64_view(sched, 1, "sched")
    confs = f64_view(confs, 2, "confs", writable=True)
    nbs = f64_view(nbs, 3, "nbs")
    mcsteps, slices, nspins = c_int(mcsteps), c_int(slices), c_int(nspins)
    if confs.shape != (nspins, slices) or nbs.shape[0] != nspins:
        raise ValueError("confs must be (nspins, slices) and nbs must have nspins rows")
    if not np.all(np.abs(confs) == 1.0):
        raise ValueError("spin configurations must hold +1/-1 values")
    if seed is None:
        seed = int(_parallel_seeds.integers(0, 2**63 - 1))
    pb = device.Problem(nbs)
    c8 = torch.from_numpy(np.ascontiguousarray(confs.T).astype(np.int8))[None].cuda()
    pb.qa_colored(c8, sched, mcsteps, c_float(temp), seed)
    confs[...] = c8[0].cpu().numpy().T.astype(np.float64)
    return None
