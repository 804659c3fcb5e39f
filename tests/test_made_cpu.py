"""Pins of the FP64 MADE oracle (CPU): finite-difference gradients, the
autoregressive property of the masks and normalisation of q."""
import itertools

import numpy as np

import made_oracle as MO


def _params(N, H1, H2, seed):
    rng = np.random.RandomState(seed)
    M1, M2, M3 = MO.masks(N, H1, H2)
    return [rng.normal(0, 0.5, M1.shape) * M1, rng.normal(0, 0.1, H1),
            rng.normal(0, 0.5, M2.shape) * M2, rng.normal(0, 0.1, H2),
            rng.normal(0, 0.5, M3.shape) * M3, rng.normal(0, 0.1, N)]


def test_gradients_match_finite_differences():
    N, H1, H2 = 6, 10, 9
    p = _params(N, H1, H2, 0)
    rng = np.random.RandomState(1)
    S = 2 * rng.randint(2, size=(7, N)) - 1
    cache, nll = MO.forward(p, S)
    g = MO.backward(p, cache)
    eps = 1e-6
    for li in range(6):
        flat = p[li].reshape(-1)
        for idx in rng.choice(flat.size, size=min(12, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            up = MO.forward(p, S)[1].mean()
            flat[idx] = old - eps
            dn = MO.forward(p, S)[1].mean()
            flat[idx] = old
            fd = (up - dn) / (2 * eps)
            assert abs(fd - g[li].reshape(-1)[idx]) < 1e-6 * max(1.0, abs(fd)), (li, idx)


def test_autoregressive_property():
    N, H1, H2 = 7, 12, 11
    p = _params(N, H1, H2, 2)
    rng = np.random.RandomState(3)
    S = 2 * rng.randint(2, size=(1, N)) - 1
    l0 = MO.forward(p, S)[0][5]
    for j in range(N):
        S2 = S.copy()
        S2[0, j] *= -1
        l1 = MO.forward(p, S2)[0][5]
        # logits d <= j never depend on spin j
        assert np.array_equal(l0[0, :j + 1], l1[0, :j + 1])


def test_q_is_normalised():
    N, H1, H2 = 8, 16, 16
    p = _params(N, H1, H2, 4)
    S = np.array(list(itertools.product([1, -1], repeat=N)))
    assert abs(np.exp(MO.logq(p, S)).sum() - 1.0) < 1e-12
