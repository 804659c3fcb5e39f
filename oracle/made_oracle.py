"""FP64 numpy restatement of the MADE network used by the learned proposal.

TEST INFRASTRUCTURE ONLY (the checker of the product's BF16 MFMA kernels).
Germain et al. 2015, natural ordering: input d has degree d+1, hidden unit k
degree 1 + k mod (N-1); masks M1[k][d] = m1(k) >= d+1, M2[k'][k] = m2(k') >=
m1(k), M3[d][k] = d+1 > m2(k).  Loss per slice: sum_d softplus(l_d) - b_d l_d
with b_d = [s_d == -1].  Pinned by central finite differences of the loss
(tests/test_made_cpu.py) and by exact enumeration of q over all 2^N slices
(normalisation).
"""
import numpy as np


def masks(N, H1, H2):
    m1 = 1 + np.arange(H1) % (N - 1)
    m2 = 1 + np.arange(H2) % (N - 1)
    d = np.arange(N)
    M1 = (m1[:, None] >= d[None, :] + 1)
    M2 = (m2[:, None] >= m1[None, :])
    M3 = (d[:, None] + 1 > m2[None, :])
    return M1.astype(np.float64), M2.astype(np.float64), M3.astype(np.float64)


def bf16(a):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    a32 = np.asarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(params, S, round_bf16=False):
    """params [w0,b0,w1,b1,w2,b2]; S int/float spins [B,N] (+-1).  Returns
    (cache, nll per row).  round_bf16 mimics the device's bf16 weights and
    activations."""
    w0, b0, w1, b1, w2, b2 = [np.asarray(p, np.float64) for p in params]
    r = bf16 if round_bf16 else (lambda a: a)
    x = np.asarray(S, np.float64)
    z1 = x @ r(w0).T + b0
    h1 = r(np.maximum(z1, 0))
    z2 = h1 @ r(w1).T + b1
    h2 = r(np.maximum(z2, 0))
    l = h2 @ r(w2).T + b2
    bits = (x < 0).astype(np.float64)
    nll = (np.logaddexp(0, l) - bits * l).sum(1)
    return (x, z1, h1, z2, h2, l, bits), nll


def backward(params, cache):
    """Gradients of the mean nll over the batch (fp64)."""
    w0, b0, w1, b1, w2, b2 = [np.asarray(p, np.float64) for p in params]
    x, z1, h1, z2, h2, l, bits = cache
    B = x.shape[0]
    dl = (1.0 / (1.0 + np.exp(-l)) - bits) / B
    gw2 = dl.T @ h2
    gb2 = dl.sum(0)
    dh2 = (dl @ w2) * (z2 > 0)
    gw1 = dh2.T @ h1
    gb1 = dh2.sum(0)
    dh1 = (dh2 @ w1) * (z1 > 0)
    gw0 = dh1.T @ x
    gb0 = dh1.sum(0)
    return [gw0, gb0, gw1, gb1, gw2, gb2]


def logq(params, S):
    return -forward(params, S)[1]
