"""Pure-Python restatement of sa.Anneal_multispin (piqmc/sa.pyx:282-405),
loop for loop, including the order in which it draws from ``rng``.  Test
infrastructure only (small sizes); the C restatement orc_sa_multispin takes
the uniforms as an array, and this one is how the tests build that array.

The one deliberate difference: the reference's unpack loop runs to nspins
inclusive (piqmc/sa.pyx:402) and touches one element past both arrays; here
only the nspins real columns are unpacked.
"""
import numpy as np


def anneal_multispin(sched, mcsteps, svec_mat, nbs, rng, record=None):
    """In place on svec_mat float64 [64, N] (bits).  If ``record`` is a list,
    (sweep temperature, permutation, [N, 64] uniforms) is appended per sweep."""
    nspins = svec_mat.shape[1]
    maxnb = nbs.shape[1]
    ediffs = np.zeros(64)
    rands = rng.rand(64)                                      # :331
    svec = np.zeros(nspins, dtype=np.uint64)
    sidx_shuff = rng.permutation(range(nspins))               # :336-337
    for si in range(svec_mat.shape[1]):                       # :339-345
        for k in range(svec_mat.shape[0]):
            svec[si] = svec[si] << np.uint64(1)
            if svec_mat[k, si]:
                svec[si] = svec[si] ^ np.uint64(1)
    for itemp in range(sched.size):
        temp = float(np.float32(sched[itemp]))                # :349 (cdef float temp)
        for step in range(mcsteps):
            used = []
            for sidx in sidx_shuff:
                for si in range(maxnb):
                    spinidx = int(nbs[sidx, si, 0])
                    jval = float(np.float32(nbs[sidx, si, 1]))  # cdef float jval
                    if spinidx == sidx:
                        w = int(svec[sidx])
                    else:
                        w = int(svec[sidx] ^ svec[spinidx])
                    for k in range(64):
                        if not (w >> (63 - k)) & 1:
                            ediffs[k] -= 2.0 * jval
                        else:
                            ediffs[k] += 2.0 * jval
                used.append(rands.copy())
                sign = np.asarray(np.exp(ediffs / temp) > rands, dtype=np.int8)
                flipmask = 1 if sign[0] else 0
                for k in range(1, 64):
                    flipmask = flipmask << 1
                    if sign[k]:
                        flipmask ^= 1
                svec[sidx] ^= np.uint64(flipmask)
                ediffs[:] = 0.0
                rands = rng.rand(64)                          # :398
            if record is not None:
                record.append((np.float32(temp), np.array(sidx_shuff), np.array(used)))
            sidx_shuff = rng.permutation(sidx_shuff)          # :400
    for sidx in range(nspins):                                # :402-405 (without the extra column)
        state = bin(int(svec[sidx]))[2:].rjust(64, '0')
        for k in range(len(state)):
            svec_mat[k, sidx] = float(state[k])
