"""Classical annealing with the parallel sweep and the device pipeline of
examples/spinglass32.py:94-137 (preanneal -> tile -> quantum anneal ->
lowest-energy slice), against the oracle's parallel sweeps."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import pm_j_problem

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _float_problem(N, rng, Z=4):
    """Random sparse couplings with fields (generic, non +-1 path)."""
    from piqmc import instances
    rows, cols, vals = [], [], []
    for i in range(N):
        for j in rng.choice(N, 2, replace=False):
            if i < j:
                rows.append(i); cols.append(j); vals.append(rng.randn())
        rows.append(i); cols.append(i); vals.append(0.3 * rng.randn())
    return instances.neighbors(N, np.array(rows), np.array(cols), np.array(vals))


@pytest.mark.parametrize("kind,C", [("pm", 8), ("pm", 128), ("float", 6), ("dense", 4)])
def test_sa_colored_matches_oracle(gpu, kind, C):
    import torch
    from piqmc import device
    rng = np.random.RandomState(C)
    if kind == "pm":
        N, nbs, _, _, _ = pm_j_problem(8, seed=C)
    else:
        N = 20
        nbs = _float_problem(N, rng)
    pb = device.Problem(nbs)
    if kind == "dense":
        pb.set_sweep_path("dense")
    sched = np.linspace(3.0, 0.05, 12)
    s0 = np.where(rng.rand(C, N) < 0.5, -1, 1).astype(np.int8)
    d = torch.from_numpy(s0.copy()).to(gpu)
    pb.sa_colored(d, sched, 2, seed=77, replica_base=5, sweep_base=3)
    ref = s0.copy()
    O.sa_colored(sched, 2, ref, nbs, pb.coloring(), pb.ncolors, 77, replica_base=5, sweep_base=3)
    assert np.array_equal(d.cpu().numpy(), ref)


def test_sa_colored_rejects_odd_chain_count(gpu):
    import torch
    from piqmc import device
    N, nbs, _, _, _ = pm_j_problem(4, seed=1)
    pb = device.Problem(nbs)
    with pytest.raises(ValueError):
        pb.sa_colored(torch.ones((3, N), dtype=torch.int8, device=gpu), [1.0], 1, 0)


@pytest.mark.parametrize("R", [6, 5])
def test_pipeline_equals_its_stages_on_the_oracle(gpu, R):
    import torch
    from piqmc import device, pipeline
    N, nbs, _, _, _ = pm_j_problem(6, seed=R)
    pb = device.Problem(nbs)
    P, T = 8, 0.1
    pre = np.linspace(3.0, 0.1, 10)
    field = np.linspace(2.0, 1e-3, 15)
    res = pipeline.anneal(pb, R, P, T, field, pre_sched=pre, seed=4)
    torch.cuda.synchronize()
    # the same stages on the oracle
    s = pipeline.random_spins(R, N, 4, "cpu").numpy().copy()
    chains = s if R % 2 == 0 else np.concatenate([s, s[:1]], 0).copy()
    O.sa_colored(pre, 1, chains, nbs, pb.coloring(), pb.ncolors, 2 * 4 + 1)
    s = chains[:R]
    E_ref = np.zeros((R, P))
    for r in range(R):
        conf = np.repeat(s[r][None, :], P, 0).copy()
        O.qa_colored(field, 1, P, T, conf, nbs, pb.coloring(), pb.ncolors, 2 * 4, replica=r)
        assert np.array_equal(res.confs[r].cpu().numpy(), conf), "replica %d" % r
        E_ref[r] = O.energy_nbs(conf, nbs)
    assert np.allclose(res.energies.cpu().numpy(), E_ref, rtol=0, atol=1e-9)
    assert np.allclose(res.best_energy.cpu().numpy(), E_ref.min(1), rtol=0, atol=1e-9)
    best_E = O.energy_nbs(res.best_conf.cpu().numpy(), nbs)
    assert np.allclose(best_E, E_ref.min(1), rtol=0, atol=1e-9)


def test_pipeline_on_the_example_instance(gpu):
    # inst_0_32x32 with the example's parameters (P=20, T=0.01, 100 field
    # steps 1.5 -> 1e-8, preanneal 3.0 -> 0.01 over 100 steps), 64 replicas.
    # No replica may go below the known ground state (examples/spinglass32.py:33),
    # and the best replica comes within 3 % of it.
    from piqmc import device, instances, pipeline
    N, rows, cols, vals = instances.read_instance(os.path.join(GOLDEN, "inst_0_32x32.txt"))
    assert N == 1024
    nbs = instances.neighbors(N, rows, cols, vals)
    pb = device.Problem(nbs)
    res = pipeline.anneal(pb, 64, 20, 0.01, np.linspace(1.5, 1e-8, 100),
                          pre_sched=np.linspace(3.0, 0.01, 100), seed=1)
    best = res.best_energy.cpu().numpy() / N
    gs = -1.55460615142578
    assert best.min() >= gs - 1e-7          # the listed value has 12 digits
    assert best.min() <= 0.97 * gs


def test_pipeline_on_the_santoro_instance(gpu):
    # the 80x80 Gaussian instance of examples/santoro80.py with its parameters
    # (P=20, T=0.01, 100 field steps 1.5 -> 1e-8, preanneal 3.0 -> 0.01 over
    # 100 steps), 64 replicas: no replica below the exact ground state
    # (santoro_80x80_answer.txt:24), the best within 4 % of it, and every
    # replica's best slice energy equal to the oracle's energy of that slice
    from piqmc import device, instances, pipeline
    N, rows, cols, vals = instances.read_instance(os.path.join(GOLDEN, "santoro_80x80.txt"))
    assert N == 6400
    nbs = instances.neighbors(N, rows, cols, vals)
    pb = device.Problem(nbs)
    res = pipeline.anneal(pb, 64, 20, 0.01, np.linspace(1.5, 1e-8, 100),
                          pre_sched=np.linspace(3.0, 0.01, 100), seed=3)
    best = res.best_energy.cpu().numpy()
    gs = -1.58051667679
    print("santoro best E/N %.6f, mean %.6f (ground state %.6f)" % (best.min() / N, best.mean() / N, gs))
    assert best.min() / N >= gs - 1e-9
    assert best.min() / N <= 0.96 * gs
    E_ref = O.energy_nbs(res.best_conf.cpu().numpy(), nbs)
    np.testing.assert_allclose(best, E_ref, rtol=0, atol=1e-8)
