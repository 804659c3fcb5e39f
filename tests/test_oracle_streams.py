"""Pins of the oracle's random streams and arithmetic primitives (CPU).

Each check compares against something outside the oracle: libc itself, numpy
itself, published known-answer vectors, or libm / closed forms.
"""
import ctypes

import numpy as np
import pytest

import oracle as O


def test_type3_matches_libc_rand():
    libc = ctypes.CDLL("libc.so.6")
    for seed in [0, 1, 2, 123, 2**31 - 1, 2**31, 2**32 - 1, 987654321, 42, 7777]:
        libc.srand(ctypes.c_uint(seed))
        t = O.Type3(seed)
        assert [libc.rand() for _ in range(1500)] == [t.next() for _ in range(1500)], seed


def test_type3_seed1_first_draws():
    # glibc's well-known first outputs after srand(1)
    t = O.Type3(1)
    assert [t.next() for _ in range(3)] == [1804289383, 846930886, 1681692777]


@pytest.mark.parametrize("seed", [0, 1, 123, 2**32 - 1, 31337])
def test_permutation_stream_matches_numpy(seed):
    # piqmc/qmc.pyx:90-91,136: rng.permutation(range(N)) then chained
    rs = np.random.RandomState(seed)
    ref = [rs.permutation(range(37))]
    for _ in range(6):
        ref.append(rs.permutation(ref[-1]))
    got = O.MT(seed).perm_stream(37, 7)
    assert np.array_equal(np.array(ref), got)


def test_permutation_stream_after_randint_use():
    # the examples draw the initial spins from the same RandomState first
    rs = np.random.RandomState(5)
    [rs.randint(2) for _ in range(10)]
    mt = O.MT(5)
    for _ in range(10):
        mt.next32()          # randint(2) consumes one masked 32-bit draw each
    ref = rs.permutation(range(64))
    assert np.array_equal(ref, mt.perm_stream(64, 1)[0])


def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32-10 (Salmon et al., SC'11)
    assert [hex(x) for x in O.philox4x32([0, 0, 0, 0], [0, 0])] == \
        ["0x6627e8d5", "0xe169c58d", "0xbc57ac4c", "0x9b00dbd8"]
    assert [hex(x) for x in O.philox4x32([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)] == \
        ["0x408f276d", "0x41c83b0e", "0xa20bc7c6", "0x6d5451fd"]
    assert [hex(x) for x in O.philox4x32([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                                         [0xA4093822, 0x299F31D0])] == \
        ["0xd16cfe09", "0x94fdcceb", "0x5001e420", "0x24126ea1"]


def test_jperp_known_values():
    # J_perp = -(P T / 2) log tanh(Gamma / (P T)), piqmc/qmc.pyx:95, stored as float.
    # Points fixed by the mathematics, not by restating the expression:
    atanh_e2 = np.arctanh(np.exp(-2.0))
    for P, T in [(20, 0.05), (64, 1 / 64), (4, 0.25), (8, 0.125)]:
        pt = P * T                                          # exactly 1 for these
        # tanh(Gamma/PT) = e^-2  =>  J_perp = PT  (the -1/2 factor, the sign, natural log)
        assert O.jperp(pt * atanh_e2, P, T) == np.float32(pt)
        # Gamma << PT: tanh x = x (1 - x^2/3 + ..), so Gamma = PT e^-10 gives 5 PT
        assert abs(O.jperp(pt * np.exp(-10.0), P, T) - 5 * pt) <= 4e-7 * pt
        # Gamma >> PT: tanh rounds to 1.0, log to 0 -> exactly zero coupling
        assert O.jperp(50.0 * pt, P, T) == 0.0
        # positive and decreasing in Gamma (ferromagnetic between slices)
        js = [O.jperp(g, P, T) for g in (1e-8, 0.01, 0.1, 1.0, 3.0)]
        assert all(a > b > 0 for a, b in zip(js, js[1:]))
    # P*T is a float product (qmc.c:2058-2075, __pyx_t_12 = slices * temp):
    # float(3 * 0.1f) = 0.30000001192.. where the double product is 0.30000000447..;
    # at Gamma = 0.01 that moves J_perp by one float ulp
    assert O.jperp(0.01, 3, 0.1) == np.float32(0.5102351903915405)
    assert O.jperp(0.01, 3, 0.1) != np.float32(0.5102351307868958)   # double P*T


def test_fexp64_accuracy():
    # The fixed-semantics acceptance exponential is within 2 ulp of libm exp.
    rng = np.random.RandomState(0)
    xs = np.concatenate([-rng.uniform(0, 745, 20000), -rng.uniform(0, 1, 20000),
                         [0.0, -1e-300, -0.5, -np.log(2), -700.0]])
    lib = O.lib()
    libm = ctypes.CDLL("libm.so.6")
    libm.exp.restype = ctypes.c_double
    libm.exp.argtypes = [ctypes.c_double]
    fexp = lib.orc_fexp64
    fexp.restype = ctypes.c_double
    fexp.argtypes = [ctypes.c_double]
    worst = 0.0
    for x in xs:
        a, b = fexp(float(x)), libm.exp(float(x))
        if b == 0.0:
            assert a < 5e-324 * 4
            continue
        worst = max(worst, abs(a - b) / np.spacing(b))
    assert worst <= 2.0, worst
    assert fexp(0.0) == 1.0
