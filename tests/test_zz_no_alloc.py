"""The allocation-free parallel sweep (piqmc_qa_colored_noalloc) against
piqmc_qa_colored, through the C ABI.

Both run the same kernels; the first takes its buffers from a fixed device
region, the second allocates them per call.  Spins and energy traces must be
identical bit for bit, call after call, with a smaller input after a larger
one (the region then holds the previous call's contents), and a call that
needs more than the region must fail with ValueError.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import pm_j_problem

pytestmark = pytest.mark.gpu


def _call(name, pb, confs, sched, M, T, seed, rbase, sbase, etrace, stream):
    from piqmc import _lib
    R, P, N = confs.shape
    sched = np.ascontiguousarray(sched, np.float64)
    _lib.call(name, pb._h, R, P, len(sched), int(M), float(T), sched.ctypes.data, confs.data_ptr(),
              int(seed), int(rbase), int(sbase), etrace.data_ptr() if etrace is not None else None,
              0, 0, C.c_void_p(stream))


def _both(gpu, nbs, R, P, S, T, seed):
    import torch
    from piqmc import device
    pb = device.Problem(nbs)
    N = nbs.shape[0]
    g = torch.Generator(device="cpu").manual_seed(seed)
    c0 = (torch.randint(0, 2, (R, P, N), generator=g, dtype=torch.int8) * 2 - 1).to(gpu)
    sched = np.linspace(3.0, 1e-8, S)
    st = torch.cuda.current_stream().cuda_stream
    a, b = c0.clone(), c0.clone()
    ea = torch.full((R, S), np.nan, dtype=torch.float64, device=gpu)
    eb = torch.full((R, S), np.nan, dtype=torch.float64, device=gpu)
    _call("piqmc_qa_colored", pb, a, sched, 1, T, seed, 3, 5, ea, st)
    _call("piqmc_qa_colored_noalloc", pb, b, sched, 1, T, seed, 3, 5, eb, st)
    torch.cuda.synchronize()
    assert not torch.equal(a, c0)                     # the sweep did something
    assert torch.equal(a, b)
    assert torch.equal(ea, eb)
    return pb


def test_noalloc_matches_allocating_entry(gpu):
    # 1024-spin +-J lattice, 256 replicas, P=64 (bit-sliced +-1 path), then
    # smaller inputs on other paths over the region the first call left:
    # a 3-D lattice with fields at P=128 (64-bit words), and a dense SK glass
    _both(gpu, pm_j_problem(32, seed=1)[1], R=256, P=64, S=40, T=1 / 64, seed=11)
    _both(gpu, pm_j_problem(4, dim=3, seed=2, field_sigma=0.1)[1], R=6, P=128, S=12, T=1 / 128, seed=12)
    from piqmc import instances
    n, r, c, v = instances.sk(64, np.random.RandomState(3))
    _both(gpu, instances.neighbors(n, r, c, v), R=4, P=16, S=10, T=1 / 16, seed=13)


def test_noalloc_rejects_oversize(gpu):
    import torch
    from piqmc import device
    pb = device.Problem(pm_j_problem(32, seed=1)[1])
    confs = torch.ones((1, 128, 1024), dtype=torch.int8, device=gpu)
    st = torch.cuda.current_stream().cuda_stream
    # 5000 replicas x 2 x 1024 spins x 8-byte words = 82 MB > 64 MiB: refused
    # before any device work, so the small buffer is never touched
    from piqmc import _lib
    sched = np.linspace(3.0, 1e-8, 4)
    with pytest.raises(ValueError, match="maximum"):
        _lib.call("piqmc_qa_colored_noalloc", pb._h, 5000, 128, 4, 1, 1 / 128.0, sched.ctypes.data,
                  confs.data_ptr(), 1, 0, 0, None, 0, 0, C.c_void_p(st))
    assert torch.equal(confs, torch.ones_like(confs))
