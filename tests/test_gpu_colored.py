"""Parallel (graph-colored, parity-alternated) sweep vs its CPU twin.

The device result must equal orc_qa_colored bit-for-bit (same phases, same
Philox4x32-10 counters, same fixed-semantics arithmetic) on both kernel paths
(bit-sliced +-1 path and per-bit general path), for every workgroup size.
"""
import numpy as np
import pytest

import oracle as O
from conftest import pm_j_problem

pytestmark = pytest.mark.gpu


def _run(gpu, nbs, P, sched, M, T, R, seed, rbase=0, sbase=0, block=0, rpw=0, coloring=None,
         path="auto"):
    import torch
    from piqmc import device
    N = nbs.shape[0]
    pb = device.Problem(nbs)
    pb.set_sweep_path(path)
    if coloring is not None:
        pb.set_coloring(coloring)
    color = pb.coloring()
    rng = np.random.RandomState(seed)
    c0 = (2 * rng.randint(2, size=(R, P, N)) - 1).astype(np.int8)
    c8 = torch.tensor(c0, device=gpu)
    et = pb.qa_colored(c8, sched, M, T, seed, replica_base=rbase, sweep_base=sbase, trace=True,
                       block=block, rpw=rpw)
    torch.cuda.synchronize()
    got, etg = c8.cpu().numpy(), et.cpu().numpy()
    ref, etr = [], []
    for r in range(R):
        sp = c0[r].copy()
        etr.append(O.qa_colored(sched, M, P, T, sp, nbs, color, int(color.max()) + 1, seed,
                                replica=rbase + r, sweep_base=sbase, trace=True))
        ref.append(sp)
    return pb, got, np.array(ref), etg, np.array(etr)


@pytest.mark.parametrize("L,P,S,T", [(8, 20, 100, 0.05), (8, 64, 40, 1 / 64), (6, 128, 10, 0.02),
                                     (4, 2, 9, 0.5), (10, 16, 30, 0.05)])
def test_colored_pm_j_bit_exact(gpu, L, P, S, T):
    N, nbs, _, _, _ = pm_j_problem(L, seed=L + P)
    sched = np.linspace(3.0, 1e-8, S)
    pb, got, ref, etg, etr = _run(gpu, nbs, P, sched, 1, T, R=5, seed=1234 + P, rbase=7, sbase=3)
    assert pb.pm_j and pb.ncolors == 2
    assert np.array_equal(got, ref)
    assert np.array_equal(etg, etr)          # +-1 energies are exact


def test_colored_3d_pm_j(gpu):
    N, nbs, _, _, _ = pm_j_problem(4, dim=3, seed=9)
    pb, got, ref, etg, etr = _run(gpu, nbs, 12, np.linspace(3.0, 1e-8, 20), 2, 0.1, R=3, seed=5)
    assert pb.pm_j and pb.zmax == 6
    assert np.array_equal(got, ref) and np.array_equal(etg, etr)


@pytest.mark.parametrize("cached", ["1", "0"])
@pytest.mark.parametrize("P,T,S", [(16, 0.05, 15), (128, 1 / 128, 12), (8, 2.0, 10), (32, 0.5, 10)])
def test_colored_fields_path(gpu, monkeypatch, P, T, S, cached):
    # 3-D +-J with Gaussian fields (the large sparse workload's structure): the
    # bit-sliced path with per-spin thresholds; T=2 and 0.5 give many
    # stochastic classes per spin (the per-bit fallback).  Both the sweep that
    # caches spins, neighbours and decision classes in registers (six
    # couplings per spin, two colors) and the general one.
    monkeypatch.setenv("PIQMC_CACHED_FIELDS", cached)
    N, nbs, _, _, _ = pm_j_problem(4, dim=3, seed=4 + P, field_sigma=0.1)
    pb, got, ref, etg, etr = _run(gpu, nbs, P, np.linspace(3.0, 1e-8, S), 1, T, R=3, seed=8 + P,
                                  rbase=2, sbase=5)
    assert pb.pm_j_fields and not pb.pm_j
    assert np.array_equal(got, ref)
    np.testing.assert_allclose(etg, etr, rtol=1e-12, atol=1e-12)


def test_colored_general_path_fields(gpu):
    # a field below 2^-27 (inexact partial sums) keeps the problem on the
    # per-bit general path, which must agree as well
    N, nbs, _, _, _ = pm_j_problem(4, dim=3, seed=4, field_sigma=0.1)
    nbs = nbs.copy()
    for e in range(nbs.shape[1]):
        if int(nbs[5, e, 0]) == 5 and nbs[5, e, 1] != 0:
            nbs[5, e, 1] = 1e-9
    pb, got, ref, etg, etr = _run(gpu, nbs, 16, np.linspace(3.0, 1e-8, 15), 1, 0.05, R=3, seed=8)
    assert not pb.pm_j and not pb.pm_j_fields
    assert np.array_equal(got, ref)
    np.testing.assert_allclose(etg, etr, rtol=1e-12, atol=1e-12)


def test_colored_general_path_frustrated(gpu):
    # odd ring + chords: non-bipartite, greedy coloring with >2 colors, float J
    from piqmc import instances
    rng = np.random.RandomState(3)
    N = 21
    rows = np.concatenate([np.arange(N), np.arange(0, N, 3)])
    cols = np.concatenate([(np.arange(N) + 1) % N, (np.arange(0, N, 3) + 7) % N])
    lo, hi = np.minimum(rows, cols), np.maximum(rows, cols)
    vals = rng.uniform(-2, 2, lo.size)
    nbs = instances.neighbors(N, lo, hi, vals)
    pb, got, ref, etg, etr = _run(gpu, nbs, 8, np.linspace(2.0, 1e-8, 25), 1, 0.1, R=4, seed=21)
    assert pb.ncolors >= 3
    assert np.array_equal(got, ref)
    np.testing.assert_allclose(etg, etr, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("field_sigma", [0.0, 0.2])
def test_colored_tile_invariance(gpu, field_sigma):
    # every workgroup size and replicas-per-workgroup choice takes identical
    # decisions (this pins every tile the autotuner may pick), on both paths
    N, nbs, _, _, _ = pm_j_problem(16 if not field_sigma else 6, seed=3, field_sigma=field_sigma)
    outs = []
    for block, rpw in [(64, 1), (128, 2), (256, 3), (512, 4), (1024, 1), (256, 5)]:
        _, got, ref, _, _ = _run(gpu, nbs, 32, np.linspace(3.0, 1e-8, 12), 1, 1 / 32, R=5, seed=77,
                                 block=block, rpw=rpw)
        assert np.array_equal(got, ref), (block, rpw)
        outs.append(got)
    assert all(np.array_equal(outs[0], o) for o in outs)


def test_colored_full_size_launch_vs_oracle(gpu):
    # the measured configuration itself: 1024-spin glass, P=64, T=1/64, 500
    # field steps, 4096 replicas in one launch with the default tiles (queue,
    # cached neighbours); the first and last replica against the CPU twin
    import torch
    from piqmc import device
    N, nbs, _, _, _ = pm_j_problem(32, seed=2024)
    pb = device.Problem(nbs)
    R, P, T = 4096, 64, 1.0 / 64
    sched = np.linspace(3.0, 1e-8, 500)
    g = torch.Generator(device="cpu").manual_seed(0)
    c0 = (torch.randint(0, 2, (R, P, N), generator=g, dtype=torch.int8) * 2 - 1)
    c8 = c0.to(gpu)
    pb.qa_colored(c8, sched, 1, T, 12345)
    torch.cuda.synchronize()
    color = pb.coloring()
    for r in (0, R - 1):
        sp = c0[r].numpy().copy()
        O.qa_colored(sched, 1, P, T, sp, nbs, color, int(color.max()) + 1, 12345, replica=r)
        assert np.array_equal(c8[r].cpu().numpy(), sp), "replica %d" % r


def test_colored_full_size_properties(gpu):
    # 1024-spin glass, P=64, T=1/64, 500 steps, 256 replicas: per-replica energy
    # trace consistent with a final slice-energy evaluation, spins stay +-1,
    # and the anneal lowers the energy to near the ground state (< -1.3/spin).
    import torch
    from piqmc import device
    N, nbs, _, _, _ = pm_j_problem(32, seed=2024)
    pb = device.Problem(nbs)
    R, P = 256, 64
    rng = np.random.RandomState(0)
    c8 = torch.tensor((2 * rng.randint(2, size=(R, P, N)) - 1).astype(np.int8), device=gpu)
    et = pb.qa_colored(c8, np.linspace(3.0, 1e-8, 500), 1, 1 / 64, 99, trace=True)
    E = pb.slice_energies(c8)
    torch.cuda.synchronize()
    assert torch.all((c8 == 1) | (c8 == -1))
    np.testing.assert_allclose(et[:, -1].cpu().numpy(), E.mean(1).cpu().numpy(), rtol=0, atol=1e-9)
    assert float(E.min(1).values.mean()) / N < -1.3


def test_colored_statistics_vs_sequential_oracle(gpu):
    """Per-field-step mean energy and ground-state hit rate of the parallel
    sweep (64 GPU replicas) within 3 sigma of the sequential fixed-semantics
    sweep (64 oracle runs, random permutation order) on a quasi-static anneal
    (20 sweeps per field step, so both orders sample the same law)."""
    import torch
    from piqmc import device
    L, P, T, S, M, nseed = 6, 16, 0.1, 30, 20, 64
    N, nbs, _, _, _ = pm_j_problem(L, seed=7)
    sched = np.linspace(3.0, 1e-8, S)
    pb = device.Problem(nbs)
    starts = []
    for s in range(nseed):
        rng = np.random.RandomState(s)
        starts.append(np.tile((2 * rng.randint(2, size=N) - 1).astype(np.int8), (P, 1)))
    c8 = torch.tensor(np.stack(starts), device=gpu)
    etg = pb.qa_colored(c8, sched, M, T, 99, trace=True).cpu().numpy()
    Eg = pb.slice_energies(c8).cpu().numpy().min(1)
    ets, Es = [], []
    for s in range(nseed):
        sp = starts[s].copy()
        _, e2 = O.qa_fixed_seq(sched, M, P, T, sp, nbs, O.MT(1000 + s).perm_stream(N, S * M),
                               O.Type3(5000 + s), trace=True)
        ets.append(e2)
        Es.append(O.energy_nbs(sp, nbs).min())
    ets, Es = np.array(ets), np.array(Es)
    se = np.sqrt(etg.var(0, ddof=1) / nseed + ets.var(0, ddof=1) / nseed)
    z = (etg.mean(0) - ets.mean(0)) / np.maximum(se, 1e-12)
    assert np.all(np.abs(z) < 3.0), z
    e_best = min(Eg.min(), Es.min())
    pg, ps = np.mean(Eg == e_best), np.mean(Es == e_best)
    pp = (pg + ps) / 2
    assert abs(pg - ps) <= 3 * np.sqrt(max(pp * (1 - pp), 1e-12) * 2 / nseed) + 1e-12, (pg, ps)


@pytest.mark.parametrize("wg", ["pipe", "pipe4", "pipe8", "1", "0"])
@pytest.mark.parametrize("tiny", [False, True])
@pytest.mark.parametrize("N,P", [(64, 16), (100, 32), (256, 128), (40, 20), (48, 8), (64, 2), (80, 64)])
def test_dense_path_sk_bit_exact(gpu, monkeypatch, N, P, tiny, wg):
    # Sherrington-Kirkpatrick couplings J_ij ~ N(0, 1/N): one color per spin,
    # the dense paths, against the CPU twin.  Such rows sum exactly in double:
    # one replica per workgroup, either pipelined (wavefront 0 decides a spin
    # while the other three or seven sum the next one's field; automatic or
    # forced count) or with the field split over four wavefronts
    # (PIQMC_DENSE_PIPE=0), or (PIQMC_DENSE_WG=0) one
    # replica per wavefront with the field split over 64/(P/2) lanes when P/2
    # divides 16.  One coupling of 1e-12 makes a row inexact and keeps the
    # row-order sum of one lane per slice.
    monkeypatch.setenv("PIQMC_DENSE_WG", "0" if wg == "0" else "1")
    monkeypatch.setenv("PIQMC_DENSE_PIPE", "1" if wg.startswith("pipe") else "0")
    monkeypatch.setenv("PIQMC_DENSE_PIPE_WAVES", {"pipe4": "4", "pipe8": "8"}.get(wg, "auto"))
    from piqmc import instances
    n, r, c, v = instances.sk(N, np.random.RandomState(N))
    if tiny:
        v = v.copy()
        v[7] = 1e-12
    nbs = instances.neighbors(n, r, c, v)
    pb, got, ref, etg, etr = _run(gpu, nbs, P, np.linspace(3.0, 1e-8, 6), 1, 0.05, R=5, seed=N + P,
                                  rbase=3)
    assert pb.ncolors == N
    assert pb.exact_rows == (not tiny)
    assert np.array_equal(got, ref)
    np.testing.assert_allclose(etg, etr, rtol=1e-12, atol=1e-10)


@pytest.mark.parametrize("waves", ["4", "8"])
@pytest.mark.parametrize("N,P,M", [(96, 48, 2), (33, 6, 3), (256, 32, 2), (200, 128, 2)])
def test_dense_pipe_sweeps_and_bases(gpu, monkeypatch, N, P, M, waves):
    # the pipelined workgroup path (three or seven summing wavefronts) over
    # several sweeps per field step, with replica and sweep bases: the spin
    # decided in one step and the field summed in the next cross sweep and
    # half-sweep boundaries
    monkeypatch.setenv("PIQMC_DENSE_WG", "1")
    monkeypatch.setenv("PIQMC_DENSE_PIPE", "1")
    monkeypatch.setenv("PIQMC_DENSE_PIPE_WAVES", waves)
    from piqmc import instances
    n, r, c, v = instances.sk(N, np.random.RandomState(N + 1))
    nbs = instances.neighbors(n, r, c, v)
    pb, got, ref, etg, etr = _run(gpu, nbs, P, np.linspace(2.0, 1e-8, 5), M, 0.1, R=3, seed=N * P,
                                  rbase=11, sbase=7)
    assert pb.exact_rows
    assert np.array_equal(got, ref)
    np.testing.assert_allclose(etg, etr, rtol=1e-12, atol=1e-10)


@pytest.mark.parametrize("field_sigma", [0.0, 0.3])
def test_dense_and_word_paths_agree(gpu, field_sigma):
    N, nbs, _, _, _ = pm_j_problem(6, seed=12, field_sigma=field_sigma)
    outs = []
    for path in ("word", "dense"):
        _, got, ref, _, _ = _run(gpu, nbs, 16, np.linspace(3.0, 1e-8, 10), 1, 0.05, R=6, seed=5,
                                 path=path)
        assert np.array_equal(got, ref), path
        outs.append(got)
    assert np.array_equal(outs[0], outs[1])


def test_batch_heterogeneous_equals_individual_runs(gpu):
    # one persistent launch over a mix of +-1 lattices (2-D and 3-D, 32- and
    # 64-bit words), a field-carrying lattice and SK glasses, with different
    # slice counts, schedule lengths and replica counts
    import torch
    from piqmc import device, instances
    jobs = []
    rng = np.random.RandomState(0)
    specs = [("pm", 8, 2, 16, 40, 5), ("pm", 16, 2, 128, 12, 3), ("pm", 4, 3, 32, 25, 4),
             ("field", 5, 2, 20, 15, 3), ("sk", 64, 0, 16, 10, 4), ("sk", 48, 0, 64, 8, 5),
             ("pm", 6, 2, 2, 30, 7)]
    for t, (kind, L, dim, P, S, R) in enumerate(specs):
        if kind == "sk":
            n, r, c, v = instances.sk(L, np.random.RandomState(t))
        else:
            n, r, c, v = instances.lattice_pm_j(L, dim, np.random.RandomState(t),
                                                 0.2 if kind == "field" else 0.0)
        pb = device.Problem(instances.neighbors(n, r, c, v))
        c0 = (2 * rng.randint(2, size=(R, P, n)) - 1).astype(np.int8)
        sched = np.linspace(3.0, 1e-8, S)
        jobs.append((pb, c0, sched, 0.05 + 0.01 * t, 100 + t, 10 * t, t))
    b = device.Batch()
    outs = []
    for pb, c0, sched, T, seed, rb, sb in jobs:
        c8 = torch.tensor(c0, device=gpu)
        et = b.add(pb, c8, sched, 1, T, seed, replica_base=rb, sweep_base=sb, trace=True)
        outs.append((c8, et))
    b.run()
    torch.cuda.synchronize()
    assert b.info()["jobs"] == len(jobs)
    for (pb, c0, sched, T, seed, rb, sb), (c8, et) in zip(jobs, outs):
        ref = torch.tensor(c0, device=gpu)
        eref = pb.qa_colored(ref, sched, 1, T, seed, replica_base=rb, sweep_base=sb, trace=True)
        torch.cuda.synchronize()
        assert torch.equal(c8, ref)
        torch.testing.assert_close(et, eref, rtol=0, atol=1e-12)


def test_batch_graph_replays_match_eager(gpu):
    # the batch launch captured once into a HIP graph; every replay (with
    # eager work and synchronisation in between) equals an eager run from the
    # same spins
    import torch
    from piqmc import device, instances
    specs = [("pm", 8, 2, 16, 20, 4), ("field", 4, 3, 32, 12, 3), ("sk", 40, 0, 16, 6, 4)]
    rng = np.random.RandomState(5)
    jobs = []
    for t, (kind, L, dim, P, S, R) in enumerate(specs):
        if kind == "sk":
            n, r, c, v = instances.sk(L, np.random.RandomState(t))
        else:
            n, r, c, v = instances.lattice_pm_j(L, dim, np.random.RandomState(t),
                                                 0.1 if kind == "field" else 0.0)
        pb = device.Problem(instances.neighbors(n, r, c, v))
        c0 = (2 * rng.randint(2, size=(R, P, n)) - 1).astype(np.int8)
        jobs.append((pb, c0, np.linspace(3.0, 1e-8, S), 0.05, 300 + t))
    bg, be = device.Batch(), device.Batch()
    cg, ce = [], []
    for pb, c0, sched, T, seed in jobs:
        a = torch.tensor(c0, device=gpu)
        e = torch.tensor(c0, device=gpu)
        bg.add(pb, a, sched, 1, T, seed)
        be.add(pb, e, sched, 1, T, seed)
        cg.append(a)
        ce.append(e)
    graph = bg.capture()
    for _ in range(3):
        graph.replay()
        torch.cuda.synchronize()
        be.run()
        torch.cuda.synchronize()
        for a, e in zip(cg, ce):
            assert torch.equal(a, e)
    assert not all(torch.equal(a, torch.tensor(c0, device=gpu)) for a, (_, c0, _, _, _) in zip(cg, jobs))


@pytest.mark.parametrize("queue", ["1", "0"])
@pytest.mark.parametrize("L,dim,P,T,S,block,rpw", [
    (32, 2, 64, 1 / 64, 40, 64, 1),     # word loop per color (no cached neighbours)
    (16, 2, 32, 1 / 32, 30, 256, 2),    # cached neighbours, degree 4
    (64, 1, 16, 0.1, 30, 128, 3),       # ring: uniform degree 2 (run-time degree)
    (4, 3, 12, 0.1, 20, 1024, 2),       # degree 6, one pass per color with idle lanes
    (8, 2, 4, 0.5, 20, 64, 1),          # many stochastic classes (inline path)
])
def test_colored_stochastic_queue(gpu, monkeypatch, queue, L, dim, P, T, S, block, rpw):
    # stochastic words drained from the per-wavefront queue, or resolved in
    # place (PIQMC_STOCHASTIC_QUEUE=0), give the oracle's result bit for bit
    monkeypatch.setenv("PIQMC_STOCHASTIC_QUEUE", queue)
    N, nbs, _, _, _ = pm_j_problem(L, dim=dim, seed=L + P)
    pb, got, ref, etg, etr = _run(gpu, nbs, P, np.linspace(3.0, 1e-8, S), 1, T, R=3, seed=31 + P,
                                  rbase=1, sbase=2, block=block, rpw=rpw)
    assert pb.pm_j
    assert np.array_equal(got, ref)
    assert np.array_equal(etg, etr)


@pytest.mark.parametrize("fixed", ["1", "0"])
@pytest.mark.parametrize("P,T,S,M", [(64, 1 / 64, 30, 1), (20, 0.05, 25, 1), (2, 0.5, 15, 1), (64, 1 / 64, 12, 3)])
def test_colored_fixed_planes(gpu, monkeypatch, fixed, P, T, S, M):
    # the 1024-spin square lattice at 256 threads and two replicas per
    # workgroup runs the fixed-plane kernel (static LDS planes of 1024 words,
    # both replicas unrolled); PIQMC_FIXED_PLANES=0 runs the general cached
    # kernel on the same tile: both equal the oracle bit for bit, with traces,
    # also with several sweeps per field step
    monkeypatch.setenv("PIQMC_FIXED_PLANES", fixed)
    N, nbs, _, _, _ = pm_j_problem(32, seed=5 + P)
    pb, got, ref, etg, etr = _run(gpu, nbs, P, np.linspace(3.0, 1e-8, S), M, T, R=4, seed=17 + P,
                                  rbase=3, sbase=5, block=256, rpw=2)
    assert pb.pm_j and pb.ncolors == 2 and N == 1024
    assert np.array_equal(got, ref)
    assert np.array_equal(etg, etr)


def test_fixed_tile_with_odd_replica_count(gpu):
    # 256 threads and two replicas per workgroup with an odd replica count: the
    # last workgroup holds one replica, so the general cached kernel runs on
    # that tile instead of the fixed-plane one; the oracle's spins either way
    N, nbs, _, _, _ = pm_j_problem(32, seed=41)
    pb, got, ref, etg, etr = _run(gpu, nbs, 64, np.linspace(3.0, 1e-8, 20), 1, 1 / 64, R=3, seed=43,
                                  rbase=1, sbase=2, block=256, rpw=2)
    assert np.array_equal(got, ref)
    assert np.array_equal(etg, etr)


@pytest.mark.parametrize("L,dim,P", [(8, 2, 64), (4, 2, 2), (4, 3, 128), (6, 2, 20), (12, 3, 66)])
@pytest.mark.parametrize("offset", [0, 1])
def test_colored_pack_roundtrip(gpu, L, dim, P, offset):
    # a call without field steps packs the spins into parity words and back:
    # the identity, through the 16-spin-group kernels (N % 16 == 0, aligned)
    # and the per-element ones (other N, or spins one byte off alignment)
    import torch
    from piqmc import device
    N, nbs, _, _, _ = pm_j_problem(L, dim=dim, seed=3)
    pb = device.Problem(nbs)
    R = 3
    rng = np.random.RandomState(P + offset)
    c0 = (2 * rng.randint(2, size=(R, P, N)) - 1).astype(np.int8)
    base = torch.empty(R * P * N + 16, dtype=torch.int8, device=gpu)
    c8 = base[offset:offset + R * P * N].view(R, P, N)
    c8.copy_(torch.from_numpy(c0))
    pb.qa_colored(c8, np.zeros(0), 1, 0.1, 1)
    torch.cuda.synchronize()
    assert np.array_equal(c8.cpu().numpy(), c0)


# ------------------------------------------------------------ autotuned tiles
@pytest.mark.parametrize("kind", ["pm", "sk"])
def test_autotuned_tile_gives_the_oracle_spins(gpu, kind):
    # the tile the autotuner picks (word path: threads and replicas per
    # workgroup; dense path: the layout of the slices over lanes and
    # wavefronts) sweeps to the CPU twin's spins
    import torch
    from piqmc import autotune, device, instances
    if kind == "pm":
        N, nbs, _, _, _ = pm_j_problem(8, seed=31)
        P, R = 16, 6
    else:
        n, r, c, v = instances.sk(64, np.random.RandomState(32))
        nbs = instances.neighbors(n, r, c, v)
        N, P, R = n, 16, 4
    pb = device.Problem(nbs)
    sched = np.linspace(3.0, 1e-8, 20)
    T = 1.0 / P
    block, rpw = autotune.tune_colored(pb, P, R, T, sched, use_cache=False)
    assert (kind == "sk") == autotune.is_dense(pb)
    rng = np.random.RandomState(5)
    c0 = (2 * rng.randint(2, size=(R, P, N)) - 1).astype(np.int8)
    c8 = torch.tensor(c0, device=gpu)
    pb.qa_colored(c8, sched, 1, T, 77, block=block, rpw=rpw)
    torch.cuda.synchronize()
    color = pb.coloring()
    for k in range(R):
        sp = c0[k].copy()
        O.qa_colored(sched, 1, P, T, sp, nbs, color, int(color.max()) + 1, 77, replica=k)
        assert np.array_equal(c8[k].cpu().numpy(), sp), (kind, block, rpw, k)


def test_dense_variants_agree_with_the_oracle(gpu):
    # every dense layout the tuner may set gives the CPU twin's spins
    import torch
    from piqmc import device, instances
    n, r, c, v = instances.sk(48, np.random.RandomState(9))
    nbs = instances.neighbors(n, r, c, v)
    pb = device.Problem(nbs)
    P, R, T = 16, 3, 1 / 16
    sched = np.linspace(3.0, 1e-8, 12)
    rng = np.random.RandomState(6)
    c0 = (2 * rng.randint(2, size=(R, P, n)) - 1).astype(np.int8)
    color = pb.coloring()
    ref = []
    for k in range(R):
        sp = c0[k].copy()
        O.qa_colored(sched, 1, P, T, sp, nbs, color, int(color.max()) + 1, 41, replica=k)
        ref.append(sp)
    for variant in device.Problem.DENSE_VARIANTS:
        pb.set_dense_variant(variant)
        c8 = torch.tensor(c0, device=gpu)
        pb.qa_colored(c8, sched, 1, T, 41)
        torch.cuda.synchronize()
        assert np.array_equal(c8.cpu().numpy(), np.array(ref)), variant


def test_autotuned_batch_equals_individual_runs(gpu):
    # the replicas per work item tune_batch picks for a heterogeneous batch
    # leave every job equal to its own qa_colored call
    import torch
    from piqmc import autotune, device, instances
    rng = np.random.RandomState(3)
    specs = [("pm", 16, 32, 30, 6), ("pm", 8, 16, 50, 5), ("sk", 64, 16, 12, 4), ("pm", 32, 64, 20, 3)]
    b, jobs = device.Batch(), []
    for t, (kind, size, P, S, R) in enumerate(specs):
        if kind == "sk":
            n, r, c, v = instances.sk(size, np.random.RandomState(t))
        else:
            n, r, c, v = instances.lattice_pm_j(size, 2, np.random.RandomState(t))
        pb = device.Problem(instances.neighbors(n, r, c, v))
        c0 = (2 * rng.randint(2, size=(R, P, n)) - 1).astype(np.int8)
        c8 = torch.tensor(c0, device=gpu)
        sched = np.linspace(3.0, 1e-8, S)
        b.add(pb, c8, sched, 1, 1.0 / P, 500 + t, replica_base=7 * t)
        jobs.append((pb, c0, c8, sched, P))
    rpw = autotune.tune_batch(b, use_cache=False)
    assert rpw in (1, 2, 4)
    b.run(rpw=rpw)
    torch.cuda.synchronize()
    for t, (pb, c0, c8, sched, P) in enumerate(jobs):
        ref = torch.tensor(c0, device=gpu)
        pb.qa_colored(ref, sched, 1, 1.0 / P, 500 + t, replica_base=7 * t)
        torch.cuda.synchronize()
        assert torch.equal(c8, ref), t


@pytest.mark.parametrize("P", [128, 32, 16])
@pytest.mark.parametrize("split", ["1", "0"])
def test_cached_fields_split_over_wavefronts(gpu, monkeypatch, split, P):
    # 256-spin SK: one replica's field tables fill a workgroup's LDS (P=128),
    # or a lone replica's chain is shared out anyway, so the P/2 slices of
    # each parity go over eight wavefronts: 8, 2 and 1 slices each, the last
    # two with update passes of two and one flipped slices
    # (PIQMC_FIELDS_SPLIT=0: one wavefront per replica); all give the CPU
    # twin's spins and energies
    from piqmc import instances
    monkeypatch.setenv("PIQMC_FIELDS_SPLIT", split)
    n, r, c, v = instances.sk(256, np.random.RandomState(4))
    nbs = instances.neighbors(n, r, c, v)
    pb, got, ref, etg, etr = _run(gpu, nbs, P, np.linspace(3.0, 1e-8, 6), 1, 1 / P, R=2, seed=17,
                                  rbase=3, sbase=1)
    assert pb.exact_rows
    assert np.array_equal(got, ref)
    np.testing.assert_allclose(etg, etr, rtol=1e-12, atol=1e-12)


def test_cached_fields_not_used_for_rows_over_256(gpu):
    # the cached-field update covers rows of up to 256 entries: a 300-spin SK
    # glass (299 couplings per row) takes another dense layout automatically
    # and gives the CPU twin's spins; asking for cached fields is refused
    import torch
    from piqmc import device, instances
    n, r, c, v = instances.sk(300, np.random.RandomState(2))
    nbs = instances.neighbors(n, r, c, v)
    pb, got, ref, etg, etr = _run(gpu, nbs, 16, np.linspace(3.0, 1e-8, 5), 1, 1 / 16, R=2, seed=23)
    assert np.array_equal(got, ref)
    pb.set_dense_variant("fields")
    c8 = torch.ones((1, 16, n), dtype=torch.int8, device=gpu)
    with pytest.raises(ValueError):
        pb.qa_colored(c8, np.linspace(3.0, 1e-8, 3), 1, 1 / 16, 1)
