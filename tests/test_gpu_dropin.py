"""The drop-in modules run the reference's own tests and match the oracle,
including the caller-visible RandomState and libc rand() states afterwards."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sps

import oracle as O
from conftest import boixo_J, init_confs, pm_j_problem

pytestmark = pytest.mark.gpu
libc = ctypes.CDLL("libc.so.6")


@pytest.fixture
def boixo():
    from piqmc import tools
    J = boixo_J()
    return J, tools.GenerateNeighbors(8, J, 4, None)


def test_reference_classicalisingenergy(gpu, boixo):
    # testing/test_boixo.py:60-70
    from piqmc import sa
    J, _ = boixo
    svecs = np.array([[-1, 1, -1, 1, 1, -1, -1, 1], [1, 1, 1, 1, 1, -1, 1, -1],
                      [1, -1, 1, 1, -1, -1, -1, -1], [1, -1, -1, 1, 1, -1, 1, 1]])
    for vec, en in zip(svecs, [4.0, -8.0, -4.0, 0.0]):
        assert sa.ClassicalIsingEnergy(vec, J) == en


def test_reference_sa(gpu, boixo):
    # testing/test_boixo.py:72-88
    from piqmc import sa
    J, nbs = boixo
    rng = np.random.RandomState(123)
    libc.srand(1)
    for _ in range(5):
        sv = np.array([2 * rng.randint(2) - 1 for _ in range(8)], dtype=np.float64)
        sa.Anneal(np.linspace(1.0, 0.01, 10), 3, sv, nbs, rng)
        assert np.sum(sv[:4]) == 4 or np.sum(sv) == -8
        assert sa.ClassicalIsingEnergy(sv, J) == -8.0


def test_reference_qmc(gpu, boixo):
    # testing/test_boixo.py:90-114
    from piqmc import qmc, sa
    J, nbs = boixo
    rng = np.random.RandomState(123)
    libc.srand(1)
    confs = np.tile(np.array([2 * rng.randint(2) - 1 for _ in range(8)], dtype=np.float64),
                    (5, 1)).T
    qmc.QuantumAnneal(np.linspace(0.5, 1e-8, 10), 3, 5, 0.01, 8, confs, nbs, rng)
    E = np.array([sa.ClassicalIsingEnergy(confs[:, k], J) for k in range(5)])
    assert np.sum(E) == -8.0 * 5
    for k in range(5):
        assert np.sum(confs[:4, k]) == 4.0 or np.sum(confs[:, k]) == -8.0


@pytest.mark.parametrize("P,T,S", [(20, 0.05, 100), (7, 0.1, 30)])
def test_quantum_anneal_dropin_matches_oracle_and_streams(gpu, P, T, S):
    from piqmc import qmc
    N, nbs, _, _, _ = pm_j_problem(8, seed=P)
    sched = np.linspace(3.0, 1e-8, S)
    rng_a, rng_b = np.random.RandomState(42), np.random.RandomState(42)
    confs_a = init_confs(N, P, rng_a)            # F-ordered like the examples
    confs_b = init_confs(N, P, rng_b).copy()
    libc.srand(2024)
    qmc.QuantumAnneal(sched, 1, P, T, N, confs_a, nbs, rng_a)
    after_libc = [libc.rand() for _ in range(3)]
    perms = [rng_b.permutation(range(N))]
    for _ in range(S):
        perms.append(rng_b.permutation(perms[-1]))
    t3 = O.Type3(2024)
    d = O.qa_compat(sched, 1, P, T, N, confs_b, nbs, np.array(perms[:S]), t3)
    assert np.array_equal(confs_a, confs_b)
    assert qmc.last_draws == d and qmc.last_near_ties == 0
    assert after_libc == [t3.next() for _ in range(3)]           # libc advanced identically
    assert rng_a.random_sample() == rng_b.random_sample()        # RandomState advanced identically


def test_anneal_dropin_matches_oracle(gpu):
    from piqmc import sa
    N, nbs, _, _, _ = pm_j_problem(8, seed=11, field_sigma=0.3)
    sched = np.linspace(3.0, 0.01, 30)
    rng_a, rng_b = np.random.RandomState(5), np.random.RandomState(5)
    sv_a = np.array([2 * rng_a.randint(2) - 1 for _ in range(N)], np.float64)
    sv_b = np.array([2 * rng_b.randint(2) - 1 for _ in range(N)], np.float64)
    libc.srand(99)
    sa.Anneal(sched, 2, sv_a, nbs, rng_a)
    perms = [rng_b.permutation(range(N))]
    for _ in range(60):
        perms.append(rng_b.permutation(perms[-1]))
    t3 = O.Type3(99)
    O.sa_compat(sched, 2, sv_b, nbs, np.array(perms[:60]), t3)
    assert np.array_equal(sv_a, sv_b)
    assert libc.rand() == t3.next()


def test_quantum_anneal_parallel_dropin(gpu):
    from piqmc import qmc, sa
    N, nbs, r, c, v = pm_j_problem(8, seed=1)
    J = sps.coo_matrix((v, (r, c)), shape=(N, N))
    rng = np.random.RandomState(0)
    confs = init_confs(N, 20, rng)
    E0 = np.mean([sa.ClassicalIsingEnergy(confs[:, k], J) for k in range(20)])
    qmc.QuantumAnneal_parallel(np.linspace(3.0, 1e-8, 100), 1, 20, 0.05, N, confs, nbs, 4, seed=3)
    E1 = [sa.ClassicalIsingEnergy(confs[:, k], J) for k in range(20)]
    assert set(np.unique(confs)) <= {-1.0, 1.0}
    assert np.mean(E1) < E0 and min(E1) / N < -1.2


def test_importing_piqmc_before_torch_keeps_the_gpu(gpu):
    # a fresh process that imports the package first and torch after it: the
    # package brings in torch's HIP runtime before loading its library, so
    # torch still sees the GPU and a device sweep runs
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from piqmc import device, instances\n"
            "import numpy as np, torch\n"
            "assert torch.cuda.is_available()\n"
            "N, r, c, v = instances.lattice_pm_j(8, 2, np.random.RandomState(1))\n"
            "pb = device.Problem(instances.neighbors(N, r, c, v))\n"
            "x = torch.ones((2, 8, N), dtype=torch.int8, device='cuda')\n"
            "pb.qa_colored(x, np.linspace(3.0, 1e-8, 5), 1, 0.1, 7)\n"
            "torch.cuda.synchronize()\n"
            "print('ok')\n") % os.path.join(root, "pathintegral-qmc_amd")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=240)
    assert out.returncode == 0 and out.stdout.strip().endswith("ok"), out.stderr[-2000:]
