"""Replay kernels vs the CPU oracle: bit-exact spins, rand() draw counts and
final libc states (qmc.QuantumAnneal, piqmc/qmc.pyx:30-136; sa.Anneal,
piqmc/sa.pyx:50-120)."""
import numpy as np
import pytest

import oracle as O
from conftest import init_confs, pm_j_problem

pytestmark = pytest.mark.gpu


def _float_instance(N, seed, fields=True, pad=2):
    """Random sparse graph with U[-2,2] couplings (the reference's generator
    range, piqmc/tools.pyx:98-130), optional fields, ragged rows (every row
    padded with at least `pad` [0, 0] entries)."""
    from piqmc import instances
    rng = np.random.RandomState(seed)
    rows, cols = [], []
    for i in range(N):
        for j in rng.choice(N, size=2, replace=False):
            if i != j:
                rows.append(min(i, j)); cols.append(max(i, j))
    rows, cols = np.array(rows), np.array(cols)
    vals = rng.uniform(-2, 2, rows.size)
    if fields:
        rows = np.concatenate([rows, np.arange(N)]); cols = np.concatenate([cols, np.arange(N)])
        vals = np.concatenate([vals, rng.uniform(-1, 1, N)])
    nbs = instances.neighbors(N, rows, cols, vals)
    return np.concatenate([nbs, np.zeros((N, pad, 2))], axis=1)


def _run_replay(gpu, nbs, P, sched, M, T, seeds, rseeds):
    import torch
    from piqmc import device
    N = nbs.shape[0]
    R, G = len(seeds), len(sched) * M
    confs0, perms = [], []
    for s in seeds:
        rng = np.random.RandomState(s)
        confs0.append(init_confs(N, P, rng))
        p = [rng.permutation(range(N))]
        for _ in range(G - 1):
            p.append(rng.permutation(p[-1]))
        perms.append(np.array(p))
    # oracle
    ref, ref_d, ref_st = [], [], []
    for r in range(R):
        c = confs0[r].copy()
        t3 = O.Type3(rseeds[r])
        ref_d.append(O.qa_compat(sched, M, P, T, N, c, nbs, perms[r], t3))
        ref.append(c)
        ref_st.append(t3.st.copy())
    # device
    pb = device.Problem(nbs)
    c8 = torch.tensor(np.stack([c.T for c in confs0]).astype(np.int8), device=gpu)
    pd = torch.tensor(np.stack(perms, axis=2).astype(np.int16), device=gpu)
    rs = torch.tensor(np.stack([O.Type3(s).st for s in rseeds]).view(np.int32), device=gpu)
    draws, ties = pb.qa_replay(c8, sched, M, T, pd, rs)
    torch.cuda.synchronize()
    got = c8.cpu().numpy()
    assert int(ties.sum()) == 0
    for r in range(R):
        assert np.array_equal(got[r], ref[r].T.astype(np.int8)), "replica %d spins" % r
        assert int(draws[r]) == ref_d[r], "replica %d draws" % r
        assert np.array_equal(rs[r].cpu().numpy().view(np.uint32), ref_st[r]), "replica %d rand" % r


@pytest.mark.parametrize("L,P,S,T", [(8, 20, 100, 0.05),    # the 64-spin +-J workload
                                     (6, 3, 9, 0.3), (4, 2, 7, 0.5), (5, 7, 6, 0.2)])
def test_qa_replay_pm_j(gpu, L, P, S, T):
    N, nbs, _, _, _ = pm_j_problem(L, seed=L * 10 + P)
    sched = np.linspace(3.0, 1e-8, S)
    _run_replay(gpu, nbs, P, sched, 1, T, seeds=list(range(100, 106)), rseeds=[1, 2, 3, 77, 0, 2**32 - 1])


def test_qa_replay_float_couplings_fields_ragged(gpu):
    nbs = _float_instance(45, 3, fields=True)
    _run_replay(gpu, nbs, 6, np.linspace(1.5, 1e-8, 12), 2, 0.01, seeds=[5, 6, 7], rseeds=[9, 10, 11])


def test_qa_replay_many_replicas(gpu):
    # 130 replicas: two full 64-lane workgroups plus a partial one
    N, nbs, _, _, _ = pm_j_problem(4, seed=1)
    _run_replay(gpu, nbs, 4, np.linspace(3.0, 1e-8, 5), 1, 0.25,
                seeds=list(range(130)), rseeds=list(range(1000, 1130)))


def test_qa_replay_wide_launch(gpu):
    # 4100 replicas: the launch shape of the 4096-replica workload (two lanes
    # per wave, two workgroups per CU) plus a partial last workgroup
    N, nbs, _, _, _ = pm_j_problem(4, seed=3)
    _run_replay(gpu, nbs, 4, np.linspace(3.0, 1e-8, 5), 1, 0.25,
                seeds=list(range(4100)), rseeds=list(range(5000, 9100)))


def test_qa_replay_full_size_subset(gpu):
    # 1024-spin +-J glass, P=64, T=1/64, 500 field steps (2 replicas vs oracle)
    N, nbs, _, _, _ = pm_j_problem(32, seed=2024)
    _run_replay(gpu, nbs, 64, np.linspace(3.0, 1e-8, 500), 1, 1.0 / 64, seeds=[1, 2], rseeds=[3, 4])


def test_sa_replay(gpu):
    import torch
    from piqmc import device
    nbs = _float_instance(70, 8, fields=True)
    N = nbs.shape[0]
    sched = np.linspace(3.0, 0.01, 20)
    M, seeds, rseeds = 2, [1, 2, 3, 4], [5, 6, 7, 8]
    sv0, perms, ref, refd = [], [], [], []
    for s, rsd in zip(seeds, rseeds):
        rng = np.random.RandomState(s)
        v = np.array([2 * rng.randint(2) - 1 for _ in range(N)], np.float64)
        p = [rng.permutation(range(N))]
        for _ in range(len(sched) * M - 1):
            p.append(rng.permutation(p[-1]))
        sv0.append(v.copy()); perms.append(np.array(p))
        refd.append(O.sa_compat(sched, M, v, nbs, np.array(p), O.Type3(rsd)))
        ref.append(v)
    pb = device.Problem(nbs)
    s8 = torch.tensor(np.stack(sv0).astype(np.int8), device=gpu)
    pd = torch.tensor(np.stack(perms, axis=2).astype(np.int16), device=gpu)
    rs = torch.tensor(np.stack([O.Type3(s).st for s in rseeds]).view(np.int32), device=gpu)
    d, t = pb.sa_replay(s8, sched, M, pd, rs)
    got = s8.cpu().numpy()
    assert int(t.sum()) == 0
    for r in range(len(seeds)):
        assert np.array_equal(got[r], ref[r].astype(np.int8))
        assert int(d[r]) == refd[r]
