"""Pins of the oracle's restatements of qmc.QuantumAnneal_dense and
sa.Anneal_dense (piqmc/qmc.pyx:141-242, piqmc/sa.pyx:126-187), CPU only.

What fixes them:
  * on matrices whose entries are float-exact and whose partial sums are
    exact, the dense functions must equal the already-pinned neighbour-table
    restatements (orc_qa_compat / orc_sa_compat) run on the same couplings:
    that checks the loop structure (D1, D2, per-spin reset, draw only on the
    Metropolis branch, the upper-triangle read);
  * hand traces for the two places where the dense functions differ from the
    neighbour-table ones: couplings read in double precision (a value that
    float rounding would collapse decides the move) and sa.Anneal_dense's
    strict ediff > 0 (piqmc/sa.pyx:181) against sa.Anneal's >= (sa.pyx:114);
  * the lower triangle is never read: garbage there changes nothing.
"""
import math

import numpy as np

import oracle as O


def _nbs_from_dense(J):
    """Neighbour table with the couplings the dense functions read: row i
    lists (si, J[min,max]) for every nonzero entry, ascending si."""
    N = J.shape[0]
    rows = []
    for i in range(N):
        r = []
        for si in range(N):
            v = J[min(i, si), max(i, si)]
            if v != 0.0:
                r.append((si, v))
        rows.append(r)
    Z = max(1, max(len(r) for r in rows))
    nbs = np.zeros((N, Z, 2))
    for i, r in enumerate(rows):
        for e, (si, v) in enumerate(r):
            nbs[i, e] = (si, v)
    return nbs


def _dyadic_upper(N, rng, density=0.5):
    """Upper-triangular J whose nonzero entries are +-2^-k with distinct
    k <= 20: any signed sum of them spans at most 21 bits, so it is exact in
    float and is never zero."""
    J = np.zeros((N, N))
    k = 1
    cells = [(a, a) for a in range(N)] + [(a, b) for a in range(N) for b in range(a + 1, N)
                                          if rng.rand() < density]
    for a, b in cells[:20]:
        J[a, b] = (1 if rng.rand() < 0.5 else -1) * 2.0 ** -k
        k += 1
    return J


def _perms(N, G, seed):
    return O.MT(seed).perm_stream(N, G)


def test_qa_dense_equals_neighbour_table_on_exact_couplings():
    rng = np.random.RandomState(3)
    for trial in range(4):
        N, P, S, M = 7, 4, 5, 2
        J = _dyadic_upper(N, rng)
        nbs = _nbs_from_dense(J)
        sched = np.linspace(2.0, 0.1, S)
        c0 = np.where(rng.rand(N, P) < 0.5, -1.0, 1.0)
        perms = _perms(N, S * M, 11 + trial)
        a, b = c0.copy(), c0.copy()
        ta, tb = O.Type3(5 + trial), O.Type3(5 + trial)
        da = O.qa_dense(sched, M, P, 0.3, N, a, J, perms, ta)
        db = O.qa_compat(sched, M, P, 0.3, N, b, nbs, perms, tb)
        assert np.array_equal(a, b) and da == db and da > 0
        assert np.array_equal(ta.st, tb.st)


def test_sa_dense_equals_neighbour_table_when_no_ediff_is_zero():
    # With no zero partial sum, > and >= decide alike.
    rng = np.random.RandomState(4)
    for trial in range(4):
        N, S, M = 9, 6, 3
        J = _dyadic_upper(N, rng)
        nbs = _nbs_from_dense(J)
        sched = np.linspace(1.5, 0.05, S)
        s0 = np.where(rng.rand(N) < 0.5, -1.0, 1.0)
        perms = _perms(N, S * M, 21 + trial)
        a, b = s0.copy(), s0.copy()
        ta, tb = O.Type3(9 + trial), O.Type3(9 + trial)
        da = O.sa_dense(sched, M, a, J, perms, ta)
        db = O.sa_compat(sched, M, b, nbs, perms, tb)
        assert np.array_equal(a, b) and da == db and da > 0


def test_trace_sa_dense_strict_accept():
    # N=1, J=[[0]]: ediff = 0.  Anneal_dense: not > 0 -> draw rand #1 and
    # exp(0)=1 > u -> flip (1 draw).  sa.Anneal would flip with no draw.
    J = np.zeros((1, 1))
    s = np.array([1.0])
    d = O.sa_dense(np.array([1.0]), 1, s, J, np.array([[0]]), O.Type3(1))
    assert d == 1 and s[0] == -1.0
    s2 = np.array([1.0])
    d2 = O.sa_compat(np.array([1.0]), 1, s2, np.array([[[0, 0.0]]]), np.array([[0]]), O.Type3(1))
    assert d2 == 0 and s2[0] == -1.0


def _double_rounding_case():
    # field h0 = 0.5, J01 = -(0.5 + 2^-50) (a double that float rounding
    # turns into -0.5), spins +1,+1, order [0, 1].
    # spin 0: e = -2*0.5 = -1, then e = (float)(-1 + -2*(J01*1)) = (float)2^-49
    #         > 0 -> flip, no draw.  (With J01 cast to float, e = 0 -> a draw.)
    # spin 1: e = (float)(-2*1*(J01*-1)) = (float)(-(1 + 2^-49)) = -1 -> draw #1,
    #         flip iff exp(-1/T) > u.
    J = np.array([[0.5, -(0.5 + 2.0 ** -50)], [123.0, 0.0]])  # J[1,0] is never read
    u = O.Type3(1).next() / 2147483647.0
    return J, u


def test_trace_sa_dense_reads_couplings_in_double():
    J, u = _double_rounding_case()
    s = np.array([1.0, 1.0])
    d = O.sa_dense(np.array([1.0]), 1, s, J, np.array([[0, 1]]), O.Type3(1))
    assert d == 1
    assert s[0] == -1.0
    assert s[1] == (-1.0 if math.exp(-1.0) > u else 1.0)


def test_trace_qa_dense_reads_couplings_in_double():
    # P=2, T=0.5, Gamma=100: P*T=1, tanh(100)=1, J_perp=0, so the Trotter
    # terms add +-0 and slice 0 follows the spatial trace above (temp 0.5).
    J, u = _double_rounding_case()
    confs = np.ones((2, 2))
    perms = np.array([[0, 1]])
    t3 = O.Type3(1)
    d = O.qa_dense(np.array([100.0]), 1, 2, 0.5, 2, confs, J, perms, t3)
    assert O.jperp(100.0, 2, 0.5) == 0.0
    assert confs[0, 0] == -1.0
    assert confs[1, 0] == (-1.0 if math.exp(-2.0) > u else 1.0)
    # slice 1 then draws for each spin whose running ediff is not > 0
    assert d >= 1


def test_dense_lower_triangle_is_never_read():
    rng = np.random.RandomState(7)
    N, P, S, M = 8, 4, 4, 2
    J = np.triu(rng.randn(N, N)) * (rng.rand(N, N) < 0.6)
    Jg = J + np.tril(rng.randn(N, N) * 1e3, -1)
    sched = np.linspace(2.0, 0.2, S)
    c0 = np.where(rng.rand(N, P) < 0.5, -1.0, 1.0)
    perms = _perms(N, S * M, 5)
    a, b = c0.copy(), c0.copy()
    da = O.qa_dense(sched, M, P, 0.4, N, a, J, perms, O.Type3(3))
    db = O.qa_dense(sched, M, P, 0.4, N, b, Jg, perms, O.Type3(3))
    assert np.array_equal(a, b) and da == db
    s1, s2 = c0[:, 0].copy(), c0[:, 0].copy()
    assert (O.sa_dense(sched, M, s1, J, perms, O.Type3(4)) ==
            O.sa_dense(sched, M, s2, Jg, perms, O.Type3(4)))
    assert np.array_equal(s1, s2)


def test_dense_strided_inputs():
    rng = np.random.RandomState(8)
    N, P = 6, 4
    J = np.triu(rng.randn(N, N))
    c0 = np.where(rng.rand(N, P) < 0.5, -1.0, 1.0)
    perms = _perms(N, 3, 2)
    sched = np.array([1.0, 0.5, 0.1])
    a = c0.copy()
    da = O.qa_dense(sched, 1, P, 0.25, N, a, J, perms, O.Type3(2))
    bF = np.asfortranarray(c0.copy())
    JF = np.asfortranarray(J)
    db = O.qa_dense(sched, 1, P, 0.25, N, bF, JF, perms, O.Type3(2))
    assert np.array_equal(a, bF) and da == db
