"""sa.Anneal_multispin drop-in and the batched packed sweep against the
restatements of piqmc/sa.pyx:282-405 (bit-exact, caller's RandomState left
in the same state)."""
import numpy as np
import pytest

import oracle as O
from conftest import pm_j_problem
from multispin_py import anneal_multispin

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("L,S,M,field", [(4, 5, 2, 0.0), (6, 8, 1, 0.3)])
def test_anneal_multispin_dropin_matches_loop_restatement(gpu, L, S, M, field):
    from piqmc import sa
    N, nbs, _, _, _ = pm_j_problem(L, seed=L, field_sigma=field)
    rng0 = np.random.RandomState(L)
    bits = (rng0.rand(64, N) < 0.5).astype(np.float64)
    sched = np.linspace(2.0, 0.1, S)
    a, b = bits.copy(), bits.copy()
    ra, rb = np.random.RandomState(9), np.random.RandomState(9)
    sa.Anneal_multispin(sched, M, a, nbs, ra)
    anneal_multispin(sched, M, b, nbs, rb)
    assert np.array_equal(a, b)
    assert sa.last_near_ties == 0
    assert ra.random_sample() == rb.random_sample()


def test_anneal_multispin_rejects_wrong_row_count(gpu):
    from piqmc import sa
    N, nbs, _, _, _ = pm_j_problem(4, seed=0)
    with pytest.raises(ValueError):
        sa.Anneal_multispin(np.array([1.0]), 1, np.zeros((32, N)), nbs, np.random.RandomState(0))


def test_batched_multispin_sets(gpu):
    import ctypes as C
    import torch
    from piqmc import _lib, device
    N, nbs, _, _, _ = pm_j_problem(8, seed=3)
    pb = device.Problem(nbs)
    R, G = 3, 7
    rng = np.random.RandomState(5)
    words = rng.randint(0, 2**64, size=(R, N), dtype=np.uint64)
    temps = np.linspace(1.5, 0.2, G).astype(np.float32)
    perms = np.stack([np.stack([rng.permutation(N) for _ in range(R)]) for _ in range(G)])  # [G][R][N]
    U = rng.rand(G, N, R, 64)
    d_bits = torch.from_numpy(words.view(np.int64).copy()).to(gpu)
    d_perms = torch.from_numpy(perms.astype(np.uint16).view(np.int16)).to(gpu)
    d_U = torch.from_numpy(U).to(gpu)
    d_ties = torch.zeros(R, dtype=torch.int32, device=gpu)
    _lib.call("piqmc_sa_multispin", pb._h, R, G, temps.ctypes.data, d_bits.data_ptr(),
              d_perms.data_ptr(), d_U.data_ptr(), d_ties.data_ptr(),
              C.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    got = d_bits.cpu().numpy().view(np.uint64)
    for r in range(R):
        w = words[r].copy()
        O.sa_multispin(temps, w, nbs, perms[:, r, :], U[:, :, r, :])
        assert np.array_equal(got[r], w), "set %d" % r
    assert int(d_ties.sum()) == 0
