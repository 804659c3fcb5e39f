"""Drop-in for piqmc.sa (piqmc/sa.pyx): classical annealing and energies.

``Anneal`` runs the replay kernel with one slice (per-spin reset, ``>=``
acceptance, float temperature per step, piqmc/sa.pyx:50-120) and reproduces
the Cython code bit-for-bit, drawing the same permutations from ``rng`` and
the same libc ``rand()`` stream.  ``Anneal_dense`` (piqmc/sa.pyx:126-187) is
the same for a dense coupling matrix.  ``Anneal_multispin`` (piqmc/sa.pyx:282-405)
runs 64 packed configurations per wavefront.  ``ClassicalIsingEnergy`` evaluates
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


#: uniforms per chunk of sweeps sent to the device at once (float64 count)
_MULTISPIN_CHUNK = 1 << 24


def Anneal_multispin(sched, mcsteps, svec_mat, nbs, rng):
    """Anneal_multispin(sched, mcsteps, svec_mat, nbs, rng)

    64 simultaneous thermal anneals, piqmc/sa.pyx:282-405.  ``svec_mat`` is
    float64 [64, nspins] of bits (nonzero = 1 = spin -1) and is overwritten
    with the final bits (0.0 / 1.0).  ``rng`` is consumed exactly as the
    reference consumes it: rand(64), permutation(range(nspins)), then per
    attempt rand(64) and per sweep permutation.  Returns None.

    Deviation: the reference's unpack loop runs to nspins inclusive
    (piqmc/sa.pyx:402) and reads/writes one element past the arrays; only the
    nspins real columns are written here.  svec_mat must have 64 rows (the
    reference's unpack writes 64 rows whatever the row count).
    """
    global last_near_ties
    import torch
    from . import device

    sched = f64_view(sched, 1, "sched")
    svec_mat = f64_view(svec_mat, 2, "svec_mat", writable=True)
    nbs = f64_view(nbs, 3, "nbs")
    mcsteps = c_int(mcsteps)
    nspins = svec_mat.shape[1]
    rands = rng.rand(64)                                   # piqmc/sa.pyx:331
    perm = rng.permutation(range(nspins))                  # :336-337
    if svec_mat.shape[0] != 64:
        raise ValueError("svec_mat must have 64 rows (one per packed configuration), got %d"
                         % svec_mat.shape[0])
    if nspins == 0:
        return None
    if nbs.shape[0] != nspins:
        raise ValueError("nbs must have one row per spin")
    bits = (svec_mat != 0)
    words = np.zeros(nspins, np.uint64)                     # :339-345
    for k in range(64):
        words = (words << np.uint64(1)) ^ bits[k].astype(np.uint64)
    S, M = sched.shape[0], max(mcsteps, 0)
    last_near_ties = 0
    if S * M > 0:
        pb = device.Problem(nbs)
        dev = torch.device("cuda", torch.cuda.current_device())
        d_bits = torch.from_numpy(words.view(np.int64).copy()).to(dev)
        d_ties = torch.zeros(1, dtype=torch.int32, device=dev)
        temps = np.repeat(sched.astype(np.float32), M)    # :349, temp = (float)sched
        per = max(1, _MULTISPIN_CHUNK // (64 * nspins))    # sweeps per device call
        for g0 in range(0, S * M, per):
            G = min(per, S * M - g0)
            perms = np.empty((G, nspins), np.uint16)
            U = np.empty((G, nspins, 64))
            for g in range(G):
                perms[g] = perm
                d = rng.rand(64 * nspins).reshape(nspins, 64)   # :398, one rand(64) per attempt
                U[g, 0] = rands
                U[g, 1:] = d[:-1]
                rands = d[-1]
                perm = rng.permutation(perm)                    # :400
            t_chunk = np.ascontiguousarray(temps[g0:g0 + G])
            d_perms = torch.from_numpy(perms.view(np.int16)).to(dev)
            d_U = torch.from_numpy(U).to(dev)
            _lib.call("piqmc_sa_multispin", pb._h, 1, G, t_chunk.ctypes.data, d_bits.data_ptr(),
                      d_perms.data_ptr(), d_U.data_ptr(), d_ties.data_ptr(),
                      C.c_void_p(torch.cuda.current_stream().cuda_stream))
            torch.cuda.synchronize()
        words = d_bits.cpu().numpy().view(np.uint64)
        last_near_ties = int(d_ties.item())
    for k in range(64):                                    # :402-405
        svec_mat[k, :] = ((words >> np.uint64(63 - k)) & np.uint64(1)).astype(np.float64)
    return None
