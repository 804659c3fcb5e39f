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


@pytest.mark.parametrize("tile", ["1", "2", "3", "4", "5"])
@pytest.mark.parametrize("M,N,K", [(64, 64, 64), (130, 70, 192), (1000, 300, 832), (1024, 1024, 1728),
                                   (2048, 512, 320), (257, 129, 64)])
def test_gemm_tiles_match_fp32_reference(gpu, monkeypatch, tile, M, N, K):
    """Every tile of the pipelined GEMM (64x64, 128x64, 128x128, 64x64 with K
    split over two or four wavefront groups), forced, including ragged edges
    in M and N and fewer K-stages than groups."""
    import torch
    from piqmc import _lib
    monkeypatch.setenv("PIQMC_GEMM_TILE", tile)
    g = torch.Generator(device="cpu").manual_seed(7 * M + N + K)
    A = torch.randn(M, K, generator=g).to(torch.bfloat16)
    B = torch.randn(N, K, generator=g).to(torch.bfloat16)
    ref = A.float() @ B.float().T
    Ad, Bd = A.view(torch.int16).to(gpu), B.view(torch.int16).to(gpu)
    C = torch.full((M, N), float("nan"), dtype=torch.float32, device=gpu)
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


@pytest.mark.parametrize("fused", [True, False])
def test_mh_move_leaves_boltzmann_law_invariant(gpu, fused):
    # frustrated 3-spin, 4-slice system at fixed field: alternate parallel
    # sweeps and MH slice moves (proposals from the fused sampler or from the
    # forward-pass sampler); the mean slice energy matches exact enumeration
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
    anneal_with_proposals(pb, m, confs, np.full(200, gamma), 1, T, seed=9, every=2, stats=stats, fused=fused)
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


# ------------------------------------------------------------ fused sampler
def _trained(N, H1, H2, seed, steps=20, lr=3e-2):
    import torch
    m = _made(N, H1, H2, seed=seed)
    rng = np.random.RandomState(seed)
    data = torch.tensor((2 * rng.randint(2, size=(256, N)) - 1).astype(np.int8), device="cuda")
    data[:, : N // 2] = 1
    for _ in range(steps):
        m.train_step(data, lr)
    return m


@pytest.mark.parametrize("N,H1,H2", [(8, 16, 16), (24, 40, 36), (3, 8, 6), (100, 150, 180), (256, 512, 512)])
@pytest.mark.parametrize("beta", [1.0, 2.5])
def test_fused_sampler_density_is_its_evaluation(gpu, N, H1, H2, beta):
    # a draw's log q_beta equals the fused evaluation of the drawn slice bit
    # for bit; at beta = 1 it is the network's log q (the FP64 oracle with the
    # device's bf16 weights and activations, to summation-order tolerance)
    import torch
    m = _trained(N, H1, H2, seed=N)
    s, lq = m.ar_sample(1000, seed=21, round_=3, beta=beta)
    assert torch.all((s == 1) | (s == -1))
    torch.testing.assert_close(m.ar_logprob(s, beta=beta), lq, rtol=0, atol=0)
    s2, _ = m.ar_sample(1000, seed=21, round_=4, beta=beta)
    assert not torch.equal(s, s2)
    if beta == 1.0:
        p = m.params()
        _, nll = MO.forward(p, s.cpu().numpy(), round_bf16=True)
        np.testing.assert_allclose(lq.cpu().numpy(), -nll, rtol=0, atol=2e-3 * N)
        np.testing.assert_allclose(m.logprob(s).cpu().numpy(), lq.cpu().numpy(), rtol=0, atol=2e-3 * N)


@pytest.mark.parametrize("beta", [1.0, 3.0])
def test_fused_sampler_is_normalized_and_samples_its_density(gpu, beta):
    # q_beta summed over all 2^8 slices is 1, and 2^15 draws follow it (chi-square)
    import torch
    N = 8
    m = _trained(N, 16, 16, seed=5, steps=30)
    allS = torch.tensor(np.array(list(itertools.product([1, -1], repeat=N)), np.int8), device=gpu)
    q = m.ar_logprob(allS, beta=beta).exp().cpu().numpy()
    assert abs(q.sum() - 1) < 1e-9
    B = 1 << 15
    s, _ = m.ar_sample(B, seed=11, round_=0, beta=beta)
    code = ((s.cpu().numpy() < 0).astype(np.int64) * (1 << np.arange(N - 1, -1, -1))).sum(1)
    counts = np.bincount(code, minlength=1 << N)
    exp = q * B
    ok = exp > 5
    chi2 = (((counts - exp) ** 2 / exp)[ok]).sum()
    dof = ok.sum() - 1
    assert chi2 < dof + 5 * np.sqrt(2 * dof), (chi2, dof)


def test_fused_sampler_rejects_unsupported_shapes(gpu):
    m = _made(160, 200, 200, seed=1)        # 130 <= N <= 193: no instantiation
    with pytest.raises(ValueError, match="fused sampler"):
        m.ar_sample(4, seed=1, round_=0)


@pytest.mark.parametrize("beta", [1.0, 2.0])
def test_worldline_move_leaves_boltzmann_law_invariant(gpu, beta):
    # frustrated 3-spin, 4-slice system at fixed field: parallel sweeps
    # alternate with world-line moves proposed by a skewed network; the mean
    # slice energy matches exact enumeration of exp(-A/T)
    import torch
    from piqmc import device
    from piqmc.made import jperp
    nbs = np.array([[[1, 1.0], [2, -1.0], [0, 0.3]],
                    [[0, 1.0], [2, 1.0], [1, -0.2]],
                    [[0, -1.0], [1, 1.0], [2, 0.1]]])
    N, P, T, gamma = 3, 4, 1.0, 1.0
    pb = device.Problem(nbs)
    m = _trained(N, 8, 8, seed=2)
    R = 4096
    confs = torch.ones((R, P, N), dtype=torch.int8, device=gpu)
    acc = 0
    for rnd in range(150):
        pb.qa_colored(confs, np.full(1, gamma), 1, T, seed=9, sweep_base=rnd)
        acc += int(m.worldline_mh(pb, confs, beta, T, seed=13, round_=rnd).sum())
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
    assert acc > 0


def test_worldline_move_decisions_2048_replicas(gpu):
    # one world-line round over 2048 replicas of the 256-spin lattice, P=32:
    # accepted replicas have every slice multiplied by f = y s^0 (slice 0
    # becomes y), rejected ones are untouched; decisions agree with the action
    # change and q ratio recomputed on the CPU where the ratio is decisive
    import torch
    from piqmc import device
    from piqmc.made import MADE
    N, nbs, _, _, _ = pm_j_problem(16, seed=1)
    pb = device.Problem(nbs)
    R, P, T = 2048, 32, 1.0 / 32
    m = MADE(N, 512, 512, seed=6)
    g = torch.Generator(device="cpu").manual_seed(3)
    confs = (torch.randint(0, 2, (R, P, N), generator=g, dtype=torch.int8) * 2 - 1).to(gpu)
    pb.qa_colored(confs, np.linspace(3.0, 1e-8, 60), 1, T, 8)
    for _ in range(10):
        m.train_step(confs[:, 0, :].contiguous(), 1e-3)
    before = confs.clone()
    acc = m.worldline_mh(pb, confs, 1.0, T, seed=17, round_=2).cpu().numpy().astype(bool)
    # the draws of that call, regenerated with the same key and round
    y, lqy = m.ar_sample(R, seed=17, round_=2, beta=1.0)
    lq0 = m.ar_logprob(before[:, 0, :].contiguous(), beta=1.0)
    b, a, yv = before.cpu().numpy(), confs.cpu().numpy(), y.cpu().numpy()
    f = yv * b[:, 0, :]
    for r in range(R):
        exp_r = b[r] * f[r][None, :] if acc[r] else b[r]
        assert np.array_equal(a[r], exp_r), r
    E0 = O.energy_nbs(b.reshape(-1, N).copy(), nbs).reshape(R, P).sum(1)
    E1 = O.energy_nbs((b * f[:, None, :]).reshape(-1, N).copy(), nbs).reshape(R, P).sum(1)
    logr = -(E1 - E0) / T + lq0.cpu().numpy() - lqy.cpu().numpy()
    must, never = logr > 1e-6, logr < -69
    assert np.all(acc[must]) and not np.any(acc[never])
    assert must.sum() + never.sum() > R // 2


def test_train_epoch_equals_its_steps(gpu):
    # the native epoch (no host synchronisation between steps) runs exactly
    # the train_step sequence: same parameters bit for bit, same mean NLL
    import torch
    from piqmc import _lib
    import ctypes as C
    N, H = 24, 40
    rng = np.random.RandomState(7)
    data = torch.tensor((2 * rng.randint(2, size=(1000, N)) - 1).astype(np.int8), device=gpu)
    perm = torch.tensor(rng.permutation(1000).astype(np.int32), device=gpu)
    a, b = _made(N, H, H, seed=9), _made(N, H, H, seed=9)
    nll = C.c_double()
    _lib.call("piqmc_made_train_epoch", a._h, 1000, data.data_ptr(), perm.data_ptr(), 128, 3e-3,
              C.byref(nll), None)
    tot = 0.0
    for t in range(1000 // 128):
        rows = data[perm[t * 128:(t + 1) * 128].long()].contiguous()
        tot += b.train_step(rows, 3e-3) * 128
    for x, y in zip(a.params(), b.params()):
        assert np.array_equal(x, y)
    assert abs(nll.value - tot / (128 * (1000 // 128))) < 1e-9 * abs(nll.value)
