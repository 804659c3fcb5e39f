"""Learned slice proposal on the GPU: BF16 MFMA GEMM vs a float32 torch
reference, network log-probabilities and gradients vs the FP64 oracle,
exactness of the sampler's density, and stationarity of the MH move."""
import itertools

import numpy as np
import pytest

import made_oracle as MO
import oracle as O
from conftest import pm_j_problem

pytestmark = pytest.mark.gpu


def _bf16_bits(t):
    import torch
    return t.to(torch.bfloat16).view(torch.int16)


@pytest.mark.parametrize("M,N,K", [(64, 64, 32), (70, 45, 100), (256, 512, 256), (1000, 300, 777),
                                   (1024, 1024, 1024), (130, 70, 192), (96, 1728, 64), (2048, 512, 320),
                                   (1024, 1024, 1728)])
def test_gemm_bf16_matches_fp32_reference(gpu, M, N, K):
    import torch
    from piqmc import _lib
    g = torch.Generator(device="cpu").manual_seed(M + N + K)
    A = torch.randn(M, K, generator=g).to(torch.bfloat16)
    B = torch.randn(N, K, generator=g).to(torch.bfloat16)
    ref = A.float() @ B.float().T                       # fp32 reference of bf16 inputs
    Ad, Bd = A.view(torch.int16).to(gpu), B.view(torch.int16).to(gpu)
    C = torch.empty(M, N, dtype=torch.float32, device=gpu)
    _lib.call("piqmc_gemm_bf16_nt", Ad.data_ptr(), Bd.data_ptr(), M, N, K, C.data_ptr(), None)
    torch.cuda.synchronize()
    torch.testing.assert_close(C.cpu(), ref, rtol=1e-4, atol=1e-3)


def _made(N=24, H1=40, H2=36, seed=3):
    from piqmc.made import MADE
    return MADE(N, H1, H2, seed=seed)


def test_logprob_and_gradients_match_oracle(gpu):
    import torch
    N, H1, H2 = 24, 40, 36
    m = _made(N, H1, H2)
    # a few Adam steps so the parameters are not at their initial values
    rng = np.random.RandomState(0)
    data = torch.tensor((2 * rng.randint(2, size=(512, N)) - 1).astype(np.int8), device=gpu)
    for _ in range(3):
        m.train_step(data, 1e-2)
    S = (2 * rng.randint(2, size=(256, N)) - 1).astype(np.int8)
    Sd = torch.tensor(S, device=gpu)
    lq = m.logprob(Sd).cpu().numpy()
    p = m.params()
    cache, nll = MO.forward(p, S, round_bf16=True)
    np.testing.assert_allclose(lq, -nll, rtol=0, atol=2e-2 * N ** 0.5)
    m.train_step(Sd, 0.0)                                 # gradients only
    gg = m.grads()
    go = MO.backward([MO.bf16(x) if i % 2 == 0 else x for i, x in enumerate(p)], cache)
    for a, b in zip(gg, go):
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert err < 3e-2, err


def test_sampler_density_is_the_evaluated_density(gpu):
    # the log q returned with each draw equals a fresh forward pass bit-for-bit,
    # and the draws follow q (chi-square over all 2^8 slices)
    import torch
    N = 8
    m = _made(N, 16, 16, seed=5)
    rng = np.random.RandomState(1)
    data = torch.tensor((2 * rng.randint(2, size=(256, N)) - 1).astype(np.int8), device=gpu)
    data[:, :4] = 1
    for _ in range(30):
        m.train_step(data, 3e-2)
    B = 1 << 15
    s, lq = m.sample(B, seed=11, round_=0)
    torch.testing.assert_close(m.logprob(s), lq, rtol=0, atol=0)
    allS = torch.tensor(np.array(list(itertools.product([1, -1], repeat=N)), np.int8), device=gpu)
    q = m.logprob(allS).exp().cpu().numpy()
    assert abs(q.sum() - 1) < 1e-4
    code = ((s.cpu().numpy() < 0).astype(np.int64) * (1 << np.arange(N - 1, -1, -1))).sum(1)
    counts = np.bincount(code, minlength=1 << N)
    exp = q * B
    ok = exp > 5
    chi2 = (((counts - exp) ** 2 / exp)[ok]).sum()
    dof = ok.sum() - 1
    assert chi2 < dof + 5 * np.sqrt(2 * dof), (chi2, dof)


def test_sampling_graph_replays_match_eager(gpu):
    # the captured sampling graph, replayed several times with eager work and
    # synchronisations in between, reproduces the eager launches every time
    import torch
    N = 16
    m = _made(N, 24, 24, seed=7)
    rng = np.random.RandomState(2)
    data = torch.tensor((2 * rng.randint(2, size=(128, N)) - 1).astype(np.int8), device=gpu)
    for _ in range(5):
        m.train_step(data, 1e-2)
    for rnd in range(4):
        m.set_graph(False)
        se, lqe = m.sample(300, seed=123, round_=rnd)
        torch.cuda.synchronize()
        m.train_step(data, 0.0)                 # unrelated eager work between replays
        m.set_graph(True)
        sg, lqg = m.sample(300, seed=123, round_=rnd)
        torch.cuda.synchronize()
        assert torch.equal(se, sg) and torch.equal(lqe, lqg), rnd
    s2, _ = m.sample(300, seed=123, round_=1)
    assert not torch.equal(s2, se)              # a different round draws differently


def test_mh_move_leaves_boltzmann_law_invariant(gpu):
    # frustrated 3-spin, 4-slice system at fixed field: alternate parallel
    # sweeps and MH slice moves; the mean slice energy matches exact enumeration
    import torch
    from piqmc import device
    from piqmc.made import anneal_with_proposals, jperp
    nbs = np.array([[[1, 1.0], [2, -1.0], [0, 0.3]],
                    [[0, 1.0], [2, 1.0], [1, -0.2]],
                    [[0, -1.0], [1, 1.0], [2, 0.1]]])
    N, P, T, gamma = 3, 4, 1.0, 4.0
    pb = device.Problem(nbs)
    m = _made(N, 8, 8, seed=2)
    rng = np.random.RandomState(4)
    data = torch.tensor((2 * rng.randint(2, size=(256, N)) - 1).astype(np.int8), device=gpu)
    data[:, 0] = 1
    for _ in range(20):
        m.train_step(data, 3e-2)        # a skewed proposal: MH must correct it
    R = 4096
    confs = torch.ones((R, P, N), dtype=torch.int8, device=gpu)
    stats = []
    anneal_with_proposals(pb, m, confs, np.full(200, gamma), 1, T, seed=9, every=2, stats=stats)
    E = pb.slice_energies(confs).mean(1).cpu().numpy()
    jp = jperp(gamma, P, T)
    S = np.array(list(itertools.product([1, -1], repeat=N * P)), np.int8).reshape(-1, P, N)
    Ek = np.stack([O.energy_nbs(S[:, k, :].copy(), nbs) for k in range(P)], 1)
    trot = sum((S[:, k, :].astype(float) * S[:, (k + 1) % P, :]).sum(1) for k in range(P))
    A = Ek.sum(1) - jp * trot
    w = np.exp(-(A - A.min()) / T)
    exact = float((w * Ek.mean(1)).sum() / w.sum())
    se = E.std(ddof=1) / np.sqrt(R)
    assert abs(E.mean() - exact) < 5 * se, (E.mean(), exact, se)
    assert np.mean(stats) > 0                           # some moves were accepted
