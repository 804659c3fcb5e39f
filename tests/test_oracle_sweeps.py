"""Pins of the oracle's sweep restatements (CPU).

Compat sweeps (qmc.QuantumAnneal / sa.Anneal with the reference's defects) are
pinned by hand-worked traces whose expected states were derived step by step
from piqmc/qmc.pyx:93-136 and piqmc/sa.pyx:93-120 (worked in the comments), by
the reference's own outcome tests (testing/test_boixo.py), and by known
ground-state energies.  Fixed-semantics sweeps (sequential and colored) are
pinned by exact enumeration of the Boltzmann law exp(-A/T) of the
(d+1)-dimensional action they must leave invariant.
"""
import itertools
import os

import numpy as np
import pytest

import oracle as O
from conftest import boixo_J

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ------------------------------------------------------------ hand traces
def test_trace_running_ediff_D2():
    # N=2, J01=+1, P=2, J_perp=0 (Gamma=1e6 -> tanh=1 -> log=0), T=1e-3.
    # slice 0 = [+1,-1], slice 1 = [+1,+1]; permutation [0,1]; srand(1).
    # slice 0: spin0 e=-2*(+1)*(1*-1)=+2 >0 -> flip; spin1 e=2+(-2*-1)*(1*-1)=0
    #          (running sum, D2) -> rand #1 = 0.84 < exp(0)=1 -> flip.
    # slice 1: spin0 e=-2 -> rand #2, exp(-2000)=0 -> stay;
    #          spin1 e=-2-2=-4 -> rand #3 -> stay.
    nbs = np.array([[[1, 1.0]], [[0, 1.0]]])
    confs = np.array([[1.0, 1.0], [-1.0, 1.0]])
    d = O.qa_compat(np.array([1e6]), 1, 2, 1e-3, 2, confs, nbs, np.array([[0, 1]]), O.Type3(1))
    assert confs.tolist() == [[-1.0, 1.0], [1.0, 1.0]]
    assert d == 3


def test_trace_fixed_trotter_slices_D1():
    # N=1 (zero field row), P=4, T=0.25, Gamma=0.01: J_perp=-0.5*log(tanh(0.01))~2.30.
    # D1: every slice couples to slices P-1=3 and 1.  Start [+1,-1,+1,-1].
    # k=0: e=+2Jp+2Jp>0 -> -1.   k=1: neighbours 3 (-1) and itself (-1):
    # e=-4Jp -> rand #1 (0.84) vs exp(-36.8) -> stay.  k=2: e=+4Jp -> -1.
    # k=3: itself and slice 1: e=-4Jp -> rand #2 -> stay.  => all -1, 2 draws.
    # (With periodic k-1,k+1 neighbours slice 1 would see e=0 and flip.)
    nbs = np.array([[[0, 0.0]]])
    confs = np.array([[1.0, -1.0, 1.0, -1.0]])
    d = O.qa_compat(np.array([0.01]), 1, 4, 0.25, 1, confs, nbs, np.array([[0]]), O.Type3(1))
    assert confs.tolist() == [[-1.0, -1.0, -1.0, -1.0]]
    assert d == 2


def test_trace_field_term():
    # N=1 field h=0.5 (self row): e=-2*s*h = +1 for s=-1 -> flip in both slices.
    nbs = np.array([[[0, 0.5]]])
    confs = np.array([[-1.0, -1.0]])
    d = O.qa_compat(np.array([1e6]), 1, 2, 1.0, 1, confs, nbs, np.array([[0]]), O.Type3(1))
    assert confs.tolist() == [[1.0, 1.0]] and d == 0


def _star_nbs():
    # spin 0 coupled to spins 1, 2, 3 with J = 0.1, 0.2, -0.3 (float32 values
    # 0.1000000015, 0.2000000030, -0.3000000119); rows padded with [0, 0]
    return np.array([[[1, 0.1], [2, 0.2], [3, -0.3]],
                     [[0, 0.1], [0, 0.0], [0, 0.0]],
                     [[0, 0.2], [0, 0.0], [0, 0.0]],
                     [[0, -0.3], [0, 0.0], [0, 0.0]]])


def test_trace_float32_running_ediff_decides():
    # Non-dyadic couplings where the reference's float32 running ediff
    # (piqmc/qmc.pyx:87,107-113; each term computed in double and stored to
    # float, qmc.c:2192-2226) takes a different decision than a double sum.
    # P=2, J_perp=0 (Gamma=1e6), T=1e-10, perm [0,1,2,3]; spin 0 = -1, others
    # +1 in both slices; srand(1).  Per slice (the slices are independent):
    #  spin0 (c=-1): e = f(0.2f) = 0.2000000030; e = f(0.2000000030 + 0.4000000060)
    #        = f(0.6000000089) = 0.6000000238; e = f(0.6000000238 - 0.6000000238)
    #        = 0 -> draw, exp(0) = 1 > u -> flip to +1.  (A double sum gives
    #        -1.49e-8 and keeps -1; dropping the (float) cast of J gives +2.4e-8
    #        and flips without a draw.)
    #  spin1: e = f(0 - 0.2000000030) -> draw, exp(-2e9) = 0 -> stay
    #  spin2: e = f(-0.2000000030 - 0.4000000060) = -0.6000000238 -> draw, stay
    #  spin3: e = f(-0.6000000238 + 0.6000000238) = 0 (running sum, D2) -> draw,
    #         flip to -1.
    # => both slices [+1, +1, +1, -1], 8 draws.
    confs = np.array([[-1.0, -1.0], [1.0, 1.0], [1.0, 1.0], [1.0, 1.0]])
    rs = O.Type3(1)
    d = O.qa_compat(np.array([1e6]), 1, 2, 1e-10, 4, confs, _star_nbs(), np.array([[0, 1, 2, 3]]), rs)
    assert confs.tolist() == [[1.0, 1.0], [1.0, 1.0], [1.0, 1.0], [-1.0, -1.0]]
    assert d == 8
    assert rs.next() == 596516649          # the 9th rand() after srand(1)


def _libc_rand():
    import ctypes
    libc = ctypes.CDLL("libc.so.6")
    return libc


def test_trace_parallel_variant_one_thread():
    # QuantumAnneal_parallel (piqmc/qmc.pyx:320-357) at nthreads=1 visits spins
    # in ascending order with one float64 ediffs[sidx] per spin, reset after
    # each slice (so per spin), J as float, D1 Trotter slices, accept on > 0
    # else exp(ediff/T) > rand()/RAND_MAX (qmc.c:4295-4345).  Same star as
    # above, P=2, J_perp=0, T=1e-10, libc srand(1).  Per slice:
    #  spin0 (c=-1): e = 0.2000000030 + 0.4000000060 - 0.6000000238 = -1.49e-8
    #        (double) -> draw, exp(-149) ~ 0 -> stay -1
    #  spin1 (c=+1, sees c0=-1): e = -2(0.1f)(-1) = +0.2 > 0 -> flip to -1
    #  spin2: e = +0.4 > 0 -> flip to -1
    #  spin3: e = -2(-0.3f)(-1) = -0.6 -> draw, stay +1
    # => both slices [-1, -1, -1, +1], 4 draws; the next rand() is the 5th.
    libc = _libc_rand()
    confs = np.array([[-1.0, -1.0], [1.0, 1.0], [1.0, 1.0], [1.0, 1.0]])
    libc.srand(1)
    O.qa_parallel(np.array([1e6]), 1, 2, 1e-10, 4, confs, _star_nbs(), 1)
    assert libc.rand() == 1957747793
    assert confs.tolist() == [[-1.0, -1.0], [-1.0, -1.0], [-1.0, -1.0], [1.0, 1.0]]


def test_trace_parallel_variant_trotter_D1():
    # The parallel variant keeps defect D1 (tidx never set, piqmc/qmc.pyx:337-345):
    # N=1 with a zero field row, P=4, T=0.25, Gamma=0.01 (J_perp ~ 2.30), start
    # [+1,-1,+1,-1].  k=0: slices 3, 1 are -1 -> e = +4 Jp > 0 -> -1.  k=1: slices
    # 3 and itself (-1) -> e = -4 Jp -> draw, exp(-36.8) -> stay.  k=2: +4 Jp ->
    # -1.  k=3: itself and slice 1 -> -4 Jp -> draw -> stay.  2 draws.  With
    # periodic neighbours k-1, k+1 slice 1 would see e = 0 at k=1 and flip.
    libc = _libc_rand()
    confs = np.array([[1.0, -1.0, 1.0, -1.0]])
    libc.srand(1)
    O.qa_parallel(np.array([0.01]), 1, 4, 0.25, 1, confs, np.array([[[0, 0.0]]]), 1)
    assert libc.rand() == 1681692777
    assert confs.tolist() == [[-1.0, -1.0, -1.0, -1.0]]


def test_trace_parallel_variant_mcsteps_and_fields():
    # Two sweeps at one field, a field row and an antiferromagnetic pair:
    # spins 0-1 with J=-1, field h0=+0.5 on spin 0; P=2, J_perp=0, T=1e-10.
    # Start slice 0 = [-1,-1], slice 1 = [+1,-1].
    # sweep 1, slice 0: spin0 e = -2(-1)(0.5f) + -2(-1)(-1 * -1) = 1 + 2 = 3 > 0 -> +1;
    #                   spin1 e = -2(-1)(-1 * +1) = -2 -> draw, stay -1.
    #          slice 1: spin0 e = -2(1)(0.5) + -2(1)(-1 * -1) = -1 - 2 = -3 -> draw, stay;
    #                   spin1 e = -2(-1)(-1 * 1) = -2 -> draw, stay.
    # sweep 2, slice 0: spin0 (+1 now) e = -1 - 2 = -3 -> draw, stay; spin1 -> draw.
    #          slice 1: as in sweep 1 -> 2 draws.  No flips; 3 + 4 = 7 draws.
    libc = _libc_rand()
    nbs = np.array([[[0, 0.5], [1, -1.0]], [[0, -1.0], [1, 0.0]]])
    confs = np.array([[-1.0, 1.0], [-1.0, -1.0]])
    libc.srand(1)
    O.qa_parallel(np.array([1e6]), 2, 2, 1e-10, 2, confs, nbs, 1)
    assert libc.rand() == 1649760492      # the 8th rand() after srand(1)
    assert confs.tolist() == [[1.0, 1.0], [-1.0, -1.0]]


def test_trace_sa_per_spin_reset_and_ge():
    # Chain 0-1-2, J=+1, rows padded with [0,0]; svec [+1,+1,-1]; perm [1,0,2].
    # spin1: e=-2*(1)+(-2)*(-1)=0 -> ">=" accepts without a draw -> [+1,-1,-1]
    # spin0: e=+2 -> flip -> [-1,-1,-1];  spin2: e=-2 -> rand #1 -> stay.
    nbs = np.array([[[1, 1.0], [0, 0.0]], [[0, 1.0], [2, 1.0]], [[1, 1.0], [0, 0.0]]])
    sv = np.array([1.0, 1.0, -1.0])
    d = O.sa_compat(np.array([1e-3]), 1, sv, nbs, np.array([[1, 0, 2]]), O.Type3(1))
    assert sv.tolist() == [-1.0, -1.0, -1.0] and d == 1


def test_compat_strided_confs_same_as_contiguous():
    rng = np.random.RandomState(3)
    N, P = 6, 5
    nbs = np.zeros((N, 2, 2))
    for i in range(N):
        nbs[i, 0] = [(i + 1) % N, rng.choice([-1.0, 1.0])]
        nbs[i, 1] = [(i - 1) % N, nbs[(i - 1) % N, 0, 1]]
    sv = 2.0 * rng.randint(2, size=N) - 1
    cf = np.tile(sv, (P, 1)).T                      # F-ordered view
    cc = np.ascontiguousarray(cf)
    perms = O.MT(9).perm_stream(N, 4)
    sched = np.linspace(2.0, 0.1, 4)
    d1 = O.qa_compat(sched, 1, P, 0.2, N, cf, nbs, perms, O.Type3(4))
    d2 = O.qa_compat(sched, 1, P, 0.2, N, cc, nbs, perms, O.Type3(4))
    assert np.array_equal(cf, cc) and d1 == d2


# ------------------------------------------------------------ reference outcome tests
def _boixo_nbs():
    from neighbors_py27 import generate_neighbors
    J = boixo_J()
    keys = list(J.keys())
    return generate_neighbors(8, keys, [J[k] for k in keys], 4)


def _boixo_energy(s):
    J = boixo_J().tocoo()
    return O.energy_coo(s, J.row, J.col, J.data)


def test_classical_energy_known_answers():
    # testing/test_boixo.py:60-70
    svecs = [[-1, 1, -1, 1, 1, -1, -1, 1], [1, 1, 1, 1, 1, -1, 1, -1],
             [1, -1, 1, 1, -1, -1, -1, -1], [1, -1, -1, 1, 1, -1, 1, 1]]
    for v, e in zip(svecs, [4.0, -8.0, -4.0, 0.0]):
        assert _boixo_energy(np.array(v, float)) == e
        assert O.energy_nbs(np.array(v, np.int8), _boixo_nbs())[0] == e


def test_boixo_sa_outcome():
    # testing/test_boixo.py:72-88 with libc rand() from srand(1)
    nbs, rng, rs = _boixo_nbs(), np.random.RandomState(123), O.Type3(1)
    for _ in range(5):
        sv = np.array([2 * rng.randint(2) - 1 for _ in range(8)], dtype=np.float64)
        sched = np.linspace(1.0, 0.01, 10)
        perms = [rng.permutation(range(8))]
        for _ in range(30):
            perms.append(rng.permutation(perms[-1]))
        O.sa_compat(sched, 3, sv, nbs, np.array(perms[:30]), rs)
        assert np.sum(sv[:4]) == 4 or np.sum(sv) == -8
        assert _boixo_energy(sv) == -8.0


def test_boixo_qmc_outcome():
    # testing/test_boixo.py:90-114 with libc rand() from srand(1)
    nbs, rng = _boixo_nbs(), np.random.RandomState(123)
    confs = np.tile(np.array([2 * rng.randint(2) - 1 for _ in range(8)], dtype=np.float64),
                    (5, 1)).T
    sched = np.linspace(0.5, 1e-8, 10)
    perms = [rng.permutation(range(8))]
    for _ in range(30):
        perms.append(rng.permutation(perms[-1]))
    O.qa_compat(sched, 3, 5, 0.01, 8, confs, nbs, np.array(perms[:30]), O.Type3(1))
    E = [_boixo_energy(confs[:, k]) for k in range(5)]
    assert np.sum(E) == -8.0 * 5
    for k in range(5):
        assert np.sum(confs[:4, k]) == 4.0 or np.sum(confs[:, k]) == -8.0


def test_instance_ground_state_energy():
    # examples/spinglass32.py:33,77 -- exact GS of the 32x32 instance
    a = np.loadtxt(os.path.join(GOLDEN, "inst_0_32x32.txt"))
    ups = np.loadtxt(os.path.join(GOLDEN, "spinglass32_gs_up.txt"), dtype=int) - 1
    gs = -np.ones(1024)
    gs[ups] = 1
    E = O.energy_coo(gs, a[:, 0].astype(np.int64) - 1, a[:, 1].astype(np.int64) - 1, a[:, 2])
    # the server's figure was computed from couplings with more digits than
    # the 12 the instance file keeps: agreement to 1e-7 per spin
    assert abs(E / 1024 - (-1.55460615142578)) < 1e-7


def test_pmj_hit_energies_fixture():
    # tests/golden/pmj_hit_energies.json: best-known energies of the +-J
    # workload instances, each with the configuration that reached it; the
    # oracle's energy of that configuration is the stored energy and no single
    # flip lowers it (a +-J lattice energy is an even integer here)
    import json
    from piqmc import instances
    with open(os.path.join(GOLDEN, "pmj_hit_energies.json")) as f:
        db = json.load(f)
    assert set(db) == {"L8_seed1", "L16_seed1", "L32_seed2024"}
    for name, d in db.items():
        N, r, c, v = instances.lattice_pm_j(d["L"], 2, np.random.RandomState(d["instance_seed"]))
        nbs = instances.neighbors(N, r, c, v)
        s = -np.ones(N, np.int8)
        s[np.array(d["config_up_1based"]) - 1] = 1
        E = float(O.energy_nbs(s[None, :].copy(), nbs)[0])
        assert E == d["energy"], name
        h = np.zeros(N)
        np.add.at(h, r, v * s[c])
        np.add.at(h, c, v * s[r])
        assert np.all(s * h >= 0), name


def test_santoro_80x80_ground_state_energy():
    # examples/ising_instances/santoro_80x80_answer.txt:24,36-256 -- exact GS of
    # the 80x80 Gaussian instance, read 1-based (the indices list the +1
    # spins; read 0-based, as examples/santoro80.py:263 does, the energy would
    # be +0.0066 per spin)
    a = np.loadtxt(os.path.join(GOLDEN, "santoro_80x80.txt"))
    ups = np.loadtxt(os.path.join(GOLDEN, "santoro_80x80_gs_up.txt"), dtype=int) - 1
    assert len(ups) == 3184                              # "spins up: 3184" (answer.txt:34)
    gs = -np.ones(6400)
    gs[ups] = 1
    r, c = a[:, 0].astype(np.int64) - 1, a[:, 1].astype(np.int64) - 1
    E = O.energy_coo(gs, r, c, a[:, 2])
    assert abs(E / 6400 - (-1.58051667679)) < 1e-9
    # a ground state: no single spin flip lowers the energy
    h = np.zeros(6400)
    np.add.at(h, r, a[:, 2] * gs[c])
    np.add.at(h, c, a[:, 2] * gs[r])
    assert np.all(gs * h >= 0)


# ------------------------------------------------------------ fixed semantics: stationarity
def _exact_mean_energy(nbs, P, T, jp):
    """<(1/P) sum_k E_k> under exp(-A/T), A = sum_k E_k - Jp sum_k,i s^k s^{k+1},
    by enumerating all 2^(N P) configurations."""
    N = nbs.shape[0]
    S = np.array(list(itertools.product([1, -1], repeat=N * P)), np.int8).reshape(-1, P, N)
    Ek = np.zeros((S.shape[0], P))
    for k in range(P):
        Ek[:, k] = O.energy_nbs(S[:, k, :].copy(), nbs)
    trot = sum((S[:, k, :].astype(float) * S[:, (k + 1) % P, :]).sum(1) for k in range(P))
    A = Ek.sum(1) - jp * trot
    w = np.exp(-(A - A.min()) / T)
    return float((w * Ek.mean(1)).sum() / w.sum())


def _small_frustrated():
    # triangle 0-1-2 (frustrated, 3 colors) with fields
    nbs = np.array([[[1, 1.0], [2, -1.0], [0, 0.3]],
                    [[0, 1.0], [2, 1.0], [1, -0.2]],
                    [[0, -1.0], [1, 1.0], [2, 0.1]]])
    return nbs, np.array([0, 1, 2], np.int32), 3


@pytest.mark.parametrize("gamma", [4.0, 6.0])
def test_fixed_semantics_stationary_law(gamma):
    nbs, color, nc = _small_frustrated()
    P, T, steps = 4, 1.0, 200000
    jp = O.jperp(gamma, P, T)
    exact = _exact_mean_energy(nbs, P, T, jp)
    sched = np.full(steps, gamma)
    # colored sweep
    sp = np.ones((P, 3), np.int8)
    et = O.qa_colored(sched, 1, P, T, sp, nbs, color, nc, seed=11, trace=True)
    # sequential sweep (random permutations, libc-style uniforms)
    sp2 = np.ones((P, 3), np.int8)
    perms = O.MT(2).perm_stream(3, steps)
    _, et2 = O.qa_fixed_seq(sched, 1, P, T, sp2, nbs, perms, O.Type3(5), trace=True)
    for trace in (et, et2):
        x = trace[1000:]
        b = x[: len(x) // 100 * 100].reshape(100, -1).mean(1)   # batch means
        se = b.std(ddof=1) / np.sqrt(len(b))
        assert abs(b.mean() - exact) < 5 * se + 1e-3, (b.mean(), exact, se)


def test_colored_rejects_improper_coloring():
    nbs, _, _ = _small_frustrated()
    sp = np.ones((4, 3), np.int8)
    with pytest.raises(ValueError):
        O.qa_colored(np.array([1.0]), 1, 4, 1.0, sp, nbs, np.array([0, 1, 1], np.int32), 2, 1)


def _exact_classical_mean_energy(nbs, T):
    N = nbs.shape[0]
    S = np.array(list(itertools.product([1, -1], repeat=N)), np.int8)
    E = O.energy_nbs(S, nbs)
    w = np.exp(-(E - E.min()) / T)
    return float((w * E).sum() / w.sum())


def test_sa_colored_stationary_law_per_step_temperature():
    # 64 chains of the frustrated triangle: sweeps at T=2.0, then at T=0.5.
    # Each phase's energy average must match the exact Boltzmann mean at the
    # temperature that phase's steps carry.
    nbs, color, nc = _small_frustrated()
    C, per_call, calls = 64, 10, 300
    sp = np.ones((C, 3), np.int8)
    for T in (2.0, 0.5):
        exact = _exact_classical_mean_energy(nbs, T)
        samples = []
        for t in range(calls):
            base = (0 if T == 2.0 else calls) * per_call + t * per_call
            O.sa_colored(np.full(per_call, T), 1, sp, nbs, color, nc, seed=5, sweep_base=base)
            if t >= 20:
                samples.append(O.energy_nbs(sp, nbs).mean())
        x = np.array(samples)
        b = x[: len(x) // 20 * 20].reshape(20, -1).mean(1)
        se = b.std(ddof=1) / np.sqrt(len(b))
        assert abs(b.mean() - exact) < 5 * se + 2e-3, (T, b.mean(), exact, se)


def test_sa_colored_is_the_parallel_sweep_with_zero_trotter_coupling():
    # Gamma = 1e3 at P*T = 1: tanh = 1, J_perp = 0; chains c = slice c%P of replica c//P
    nbs, color, nc = _small_frustrated()
    C, T = 8, 0.7
    rng = np.random.RandomState(1)
    sp = np.where(rng.rand(C, 3) < 0.5, -1, 1).astype(np.int8)
    a = sp.copy()
    O.sa_colored(np.full(5, T), 2, a, nbs, color, nc, seed=9, replica_base=3, sweep_base=4)
    b = sp.copy()
    assert O.jperp(1e3 * T * C, C, T) == 0.0
    # one replica of P=C slices; qa_colored's temp is the same T at every step
    O.qa_colored(np.full(5, 1e3 * T * C), 2, C, T, b, nbs, color, nc, seed=9, replica=3,
                 sweep_base=4)
    assert np.array_equal(a, b)
