"""Pins of the Anneal_multispin restatements (piqmc/sa.pyx:282-405), CPU.

* Decomposition: the 64 packed configurations are 64 independent +-1
  Metropolis chains; each one, simulated on its own with +-1 spins (no bit
  tricks) and fed the same permutation and its own uniforms, must end where
  the packed restatement leaves its bit.  This fixes the bit order (63-k),
  the XOR/field branches and the sign of ediff.
* The C restatement equals the pure-Python loop restatement given the
  uniforms that loop drew, and the packing helpers round-trip.
* Hand trace on one spin.
"""
import numpy as np

import oracle as O
from multispin_py import anneal_multispin


def _problem(N, rng, Z=3):
    nbs = np.zeros((N, Z, 2))
    for i in range(N):
        nbs[i, 0] = (i, 0.25 * rng.randn())                 # field
        for z in range(1, Z):
            j = rng.randint(N)
            nbs[i, z] = (j, rng.choice([-1.0, 1.0, 0.5]))
    # symmetrise: make every coupling appear in both rows with the same value
    for i in range(N):
        for z in range(1, Z):
            j = int(nbs[i, z, 0])
            if j == i:
                continue
            if not any(int(nbs[j, y, 0]) == i for y in range(Z)):
                nbs[i, z] = (i, 0.0)
    return nbs


def _chains(bits, nbs, temps, perms, U):
    """64 independent +-1 chains (configuration k = bits row k)."""
    S = np.where(bits != 0, -1.0, 1.0)                       # [64, N]
    for g, T in enumerate(temps):
        for t, i in enumerate(perms[g]):
            e = np.zeros(64)
            for z in range(nbs.shape[1]):
                j, J = int(nbs[i, z, 0]), float(np.float32(nbs[i, z, 1]))
                if j == i:
                    e = e + (-2.0 * S[:, i]) * J
                else:
                    e = e + (-2.0 * S[:, i] * S[:, j]) * J
            acc = np.exp(e / float(T)) > U[g, t]
            S[acc, i] *= -1.0
    return (S < 0).astype(np.float64)


def test_multispin_is_64_independent_chains():
    rng = np.random.RandomState(0)
    N, G = 12, 6
    nbs = _problem(N, rng)
    bits = (rng.rand(64, N) < 0.5).astype(np.float64)
    temps = np.linspace(2.0, 0.2, G).astype(np.float32)
    perms = np.array([rng.permutation(N) for _ in range(G)])
    U = rng.rand(G, N, 64)
    w = O.pack_multispin(bits)
    O.sa_multispin(temps, w, nbs, perms, U)
    got = O.unpack_multispin(w)
    assert np.array_equal(got, _chains(bits, nbs, temps, perms, U))
    assert 0 < np.abs(got - bits).sum()                      # something moved


def test_python_loop_and_c_restatements_agree():
    rng = np.random.RandomState(1)
    N = 9
    nbs = _problem(N, rng)
    bits = (rng.rand(64, N) < 0.5).astype(np.float64)
    sched, M = np.array([1.5, 0.8, 0.3]), 2
    a = bits.copy()
    rec = []
    r1 = np.random.RandomState(42)
    anneal_multispin(sched, M, a, nbs, r1, record=rec)
    temps = np.array([t for t, _, _ in rec], np.float32)
    perms = np.array([p for _, p, _ in rec])
    U = np.array([u for _, _, u in rec])
    assert U.shape == (len(sched) * M, N, 64)
    w = O.pack_multispin(bits)
    O.sa_multispin(temps, w, nbs, perms, U)
    assert np.array_equal(O.unpack_multispin(w), a)
    # stream order: rand(64), permutation, then rand(64) per attempt and a
    # permutation per sweep -- replay it by hand with a fresh generator
    r2 = np.random.RandomState(42)
    first = r2.rand(64)
    p0 = r2.permutation(range(N))
    assert np.array_equal(U[0, 0], first) and np.array_equal(perms[0], p0)
    draws = [r2.rand(64) for _ in range(N)]
    assert all(np.array_equal(U[0, t], draws[t - 1]) for t in range(1, N))
    assert np.array_equal(perms[1], r2.permutation(p0))
    assert np.array_equal(U[1, 0], draws[-1])


def test_pack_round_trip():
    rng = np.random.RandomState(2)
    bits = (rng.rand(64, 20) < 0.5).astype(np.float64)
    w = O.pack_multispin(bits)
    assert np.array_equal(O.unpack_multispin(w), bits)
    assert int(w[0]) >> 63 == int(bits[0, 0])               # row 0 is the top bit


def test_trace_single_spin_field():
    # N=1, field h=0.5: configuration with bit 0 (s=+1) has ediff=-1, bit 1 has +1.
    # At T=1: bit-1 configurations flip iff e^1 > u (always); bit-0 iff e^-1 > u.
    nbs = np.array([[[0, 0.5]]])
    bits = np.zeros((64, 1))
    bits[1::2] = 1.0
    U = np.full((1, 1, 64), 0.5)
    U[0, 0, 0] = 0.2                                        # e^-1 = 0.368 > 0.2 -> flips
    w = O.pack_multispin(bits)
    O.sa_multispin(np.array([1.0], np.float32), w, nbs, np.array([[0]]), U)
    out = O.unpack_multispin(w)[:, 0]
    exp_out = 1.0 - bits[:, 0]                               # bit-1 rows flip to 0
    exp_out[0::2] = 0.0                                      # bit-0 rows stay (0.368 < 0.5) ...
    exp_out[0] = 1.0                                         # ... except row 0 (u = 0.2)
    assert np.array_equal(out, exp_out)
