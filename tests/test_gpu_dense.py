"""qmc.QuantumAnneal_dense / sa.Anneal_dense drop-ins and replay on a problem
built from a dense matrix: bit-exact against the oracle's restatements of
piqmc/qmc.pyx:141-242 and piqmc/sa.pyx:126-187, including the caller-visible
RandomState and libc rand() states afterwards."""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
libc = ctypes.CDLL("libc.so.6")


def _sk(N, rng, field=0.1, lower_garbage=False):
    J = np.triu(rng.randn(N, N) / np.sqrt(N), 1)
    J[np.diag_indices(N)] = field * rng.randn(N)
    if lower_garbage:
        J += np.tril(rng.randn(N, N) * 100.0, -1)
    return J


def _perm_list(rng, N, G):
    perms = [rng.permutation(range(N))]
    for _ in range(G):
        perms.append(rng.permutation(perms[-1]))
    return np.array(perms[:G])


@pytest.mark.parametrize("N,P,T,S,M", [(24, 4, 0.2, 12, 2), (40, 8, 0.05, 20, 1), (9, 2, 1.0, 5, 3)])
def test_quantum_anneal_dense_dropin_matches_oracle(gpu, N, P, T, S, M):
    from piqmc import qmc
    rng_j = np.random.RandomState(N + P)
    J = _sk(N, rng_j, lower_garbage=True)
    sched = np.linspace(3.0, 1e-3, S)
    rng_a, rng_b = np.random.RandomState(7), np.random.RandomState(7)
    c0 = np.where(rng_j.rand(N, P) < 0.5, -1.0, 1.0)
    confs_a = np.asfortranarray(c0.copy())
    confs_b = c0.copy()
    libc.srand(99)
    qmc.QuantumAnneal_dense(sched, M, P, T, N, confs_a, J, rng_a)
    after = [libc.rand() for _ in range(3)]
    perms = _perm_list(rng_b, N, S * M)
    t3 = O.Type3(99)
    d = O.qa_dense(sched, M, P, T, N, confs_b, J, perms, t3)
    assert np.array_equal(confs_a, confs_b)
    assert qmc.last_draws == d and d > 0 and qmc.last_near_ties == 0
    assert after == [t3.next() for _ in range(3)]
    assert rng_a.random_sample() == rng_b.random_sample()


@pytest.mark.parametrize("N,S,M", [(30, 15, 2), (64, 10, 1)])
def test_anneal_dense_dropin_matches_oracle(gpu, N, S, M):
    from piqmc import sa
    rng_j = np.random.RandomState(N)
    J = _sk(N, rng_j, field=0.3)
    sched = np.linspace(2.0, 0.01, S)
    rng_a, rng_b = np.random.RandomState(3), np.random.RandomState(3)
    s0 = np.where(rng_j.rand(N) < 0.5, -1.0, 1.0)
    a, b = s0.copy(), s0.copy()
    libc.srand(5)
    sa.Anneal_dense(sched, M, a, J, rng_a)
    perms = _perm_list(rng_b, N, S * M)
    t3 = O.Type3(5)
    d = O.sa_dense(sched, M, b, J, perms, t3)
    assert np.array_equal(a, b)
    assert sa.last_draws == d and sa.last_near_ties == 0
    assert libc.rand() == t3.next()


def test_anneal_dense_strict_accept_on_gpu(gpu):
    # the oracle's hand trace: a zero ediff draws under Anneal_dense
    from piqmc import sa
    J = np.zeros((1, 1))
    s = np.array([1.0])
    libc.srand(1)
    sa.Anneal_dense(np.array([1.0]), 1, s, J, np.random.RandomState(0))
    assert sa.last_draws == 1 and s[0] == -1.0


def test_dense_double_rounding_trace_on_gpu(gpu):
    # J01 = -(0.5 + 2^-50): read in double it decides spin 0's move without a draw
    from piqmc import sa
    J = np.array([[0.5, -(0.5 + 2.0 ** -50)], [123.0, 0.0]])
    rng = np.random.RandomState(0)
    while True:                          # a seed whose first permutation is [0, 1]
        st = rng.get_state()
        if list(rng.permutation(range(2))) == [0, 1]:
            rng.set_state(st)
            break
    s = np.array([1.0, 1.0])
    libc.srand(1)
    sa.Anneal_dense(np.array([1.0]), 1, s, J, rng)
    assert sa.last_draws == 1 and s[0] == -1.0


def test_batched_replay_on_dense_problem(gpu):
    import torch
    from piqmc import device
    N, P, T, S, R = 32, 4, 0.1, 8, 5
    rng = np.random.RandomState(11)
    J = _sk(N, rng)
    pb = device.Problem.from_dense(J)
    sched = np.linspace(2.5, 1e-3, S)
    seeds = np.arange(R, dtype=np.uint32) + 40
    perms = device.perm_streams(seeds, N, S)                         # [S, N, R]
    c0 = np.where(rng.rand(R, P, N) < 0.5, -1, 1).astype(np.int8)
    rs = device.rand_states(np.arange(R) + 300)
    c8 = torch.from_numpy(c0.copy()).to(gpu)
    pd = torch.from_numpy(perms.view(np.int16)).to(gpu)
    rd = torch.from_numpy(rs.view(np.int32).copy()).to(gpu)
    draws, ties = pb.qa_replay(c8, sched, 1, T, pd, rd)
    torch.cuda.synchronize()
    got = c8.cpu().numpy()
    for r in range(R):
        conf = c0[r].T.astype(np.float64).copy()
        t3 = O.Type3(state=rs[r])
        d = O.qa_dense(sched, 1, P, T, N, conf, J, perms[:, :, r].astype(np.int64), t3)
        assert np.array_equal(got[r], conf.T.astype(np.int8)), "replica %d" % r
        assert int(draws[r]) == d
        assert np.array_equal(rd[r].cpu().numpy().view(np.uint32), t3.st)
    assert int(ties.sum()) == 0


def test_dense_problem_rejects_nonfinite(gpu):
    from piqmc import device
    J = np.zeros((3, 3))
    J[0, 2] = np.inf
    with pytest.raises(ValueError):
        device.Problem.from_dense(J)
