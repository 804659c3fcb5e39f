"""ctypes front end of the CPU oracle (oracle/piqmc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "liborc.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
        i64p = np.ctypeslib.ndpointer(np.int64, flags="C")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
        i8p = np.ctypeslib.ndpointer(np.int8, flags="C")
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
        L.orc_rand_next.argtypes = [u32p]
        L.orc_rand_next.restype = C.c_uint32
        L.orc_rand_seed.argtypes = [C.c_uint32, u32p]
        L.orc_mt_seed.argtypes = [C.c_uint32, u32p]
        L.orc_mt_next32.argtypes = [u32p]
        L.orc_mt_next32.restype = C.c_uint32
        L.orc_mt_shuffle_i64.argtypes = [u32p, i64p, C.c_int64]
        L.orc_perm_stream.argtypes = [u32p, C.c_int, C.c_int, i64p]
        L.orc_philox4x32.argtypes = [u32p, u32p, u32p]
        L.orc_jperp.argtypes = [C.c_double, C.c_int, C.c_float]
        L.orc_jperp.restype = C.c_float
        L.orc_qa_compat.argtypes = [f64p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int,
                                    C.c_void_p, C.c_long, C.c_long, f64p, C.c_int,
                                    i64p, u32p, C.POINTER(C.c_uint64)]
        L.orc_sa_compat.argtypes = [f64p, C.c_int, C.c_int, C.c_void_p, C.c_long, C.c_int,
                                    f64p, C.c_int, i64p, u32p, C.POINTER(C.c_uint64)]
        L.orc_qa_dense.argtypes = [f64p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int,
                                   C.c_void_p, C.c_long, C.c_long, C.c_void_p, C.c_long, C.c_long,
                                   i64p, u32p, C.POINTER(C.c_uint64)]
        L.orc_sa_dense.argtypes = [f64p, C.c_int, C.c_int, C.c_void_p, C.c_long, C.c_int,
                                   C.c_void_p, C.c_long, C.c_long, i64p, u32p,
                                   C.POINTER(C.c_uint64)]
        L.orc_qa_parallel.argtypes = [f64p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int,
                                      C.c_void_p, C.c_long, C.c_long, f64p, C.c_int, C.c_int]
        L.orc_qa_fixed_seq.argtypes = [f64p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int,
                                       i8p, f64p, C.c_int, i64p, u32p,
                                       C.POINTER(C.c_uint64), C.c_void_p]
        L.orc_qa_colored.argtypes = [f64p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int,
                                     i8p, f64p, C.c_int, i32p, C.c_int, C.c_uint64,
                                     C.c_uint32, C.c_uint32, C.c_void_p]
        L.orc_qa_colored_stats.argtypes = [f64p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int,
                                           i8p, f64p, C.c_int, i32p, C.c_int, C.c_uint64,
                                           C.c_uint32, np.ctypeslib.ndpointer(np.uint64, flags="C")]
        L.orc_sa_colored.argtypes = [f64p, C.c_int, C.c_int, C.c_int, C.c_int, i8p, f64p,
                                     C.c_int, i32p, C.c_int, C.c_uint64, C.c_uint32, C.c_uint32]
        L.orc_sa_multispin.argtypes = [np.ctypeslib.ndpointer(np.float32, flags="C"), C.c_int,
                                       C.c_int, np.ctypeslib.ndpointer(np.uint64, flags="C"),
                                       f64p, C.c_int, i64p, f64p]
        L.orc_sa_multispin.restype = C.c_int
        L.orc_energy_nbs.argtypes = [i8p, C.c_int, C.c_int, f64p, C.c_int, f64p]
        L.orc_energy_coo.argtypes = [f64p, C.c_int, C.c_long, i64p, i64p, f64p]
        L.orc_energy_coo.restype = C.c_double
        for name in ("orc_qa_compat", "orc_sa_compat", "orc_qa_dense", "orc_sa_dense",
                     "orc_qa_parallel",
                     "orc_qa_fixed_seq", "orc_qa_colored", "orc_qa_colored_stats", "orc_sa_colored"):
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


# ---------------------------------------------------------------- streams
class Type3:
    """glibc random() TYPE_3 generator state (31 words + position)."""

    def __init__(self, seed=1, state=None):
        if state is not None:
            self.st = np.array(state, dtype=np.uint32).copy()
        else:
            self.st = np.zeros(32, np.uint32)
            lib().orc_rand_seed(seed & 0xFFFFFFFF, self.st)

    def next(self):
        return lib().orc_rand_next(self.st)

    def draws(self, n):
        return np.array([self.next() for _ in range(n)], dtype=np.uint32)


class MT:
    """numpy legacy RandomState(seed) MT19937 state (624 words + position)."""

    def __init__(self, seed):
        self.st = np.zeros(625, np.uint32)
        lib().orc_mt_seed(seed & 0xFFFFFFFF, self.st)

    def next32(self):
        return lib().orc_mt_next32(self.st)

    def perm_stream(self, n, g):
        out = np.empty((g, n), np.int64)
        lib().orc_perm_stream(self.st, n, g, out)
        return out


def philox4x32(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def jperp(gamma, P, temp):
    return lib().orc_jperp(float(gamma), int(P), float(temp))


# ---------------------------------------------------------------- sweeps
def _strides(a):
    return a.strides[0] // a.itemsize, a.strides[1] // a.itemsize


def qa_compat(sched, mcsteps, P, temp, N, confs, nbs, perms, rstate):
    """qmc.QuantumAnneal restatement; confs (f64 [N,P], any strides) in place.
    perms: int64 [S*M (+1), N]. rstate: Type3 (advanced). Returns draws."""
    sched = np.ascontiguousarray(sched, np.float64)
    nbs = np.ascontiguousarray(nbs, np.float64)
    perms = np.ascontiguousarray(perms, np.int64)
    assert confs.dtype == np.float64
    cs0, cs1 = _strides(confs)
    d = C.c_uint64(0)
    rc = lib().orc_qa_compat(sched, len(sched), int(mcsteps), int(P), float(temp), int(N),
                             confs.ctypes.data, cs0, cs1, nbs, nbs.shape[1], perms,
                             rstate.st, C.byref(d))
    if rc:
        raise ValueError("orc_qa_compat rc=%d" % rc)
    return d.value


def sa_compat(sched, mcsteps, svec, nbs, perms, rstate):
    sched = np.ascontiguousarray(sched, np.float64)
    nbs = np.ascontiguousarray(nbs, np.float64)
    perms = np.ascontiguousarray(perms, np.int64)
    assert svec.dtype == np.float64
    d = C.c_uint64(0)
    rc = lib().orc_sa_compat(sched, len(sched), int(mcsteps), svec.ctypes.data,
                             svec.strides[0] // 8, svec.shape[0], nbs, nbs.shape[1],
                             perms, rstate.st, C.byref(d))
    if rc:
        raise ValueError("orc_sa_compat rc=%d" % rc)
    return d.value


def qa_dense(sched, mcsteps, P, temp, N, confs, J, perms, rstate):
    """qmc.QuantumAnneal_dense restatement; confs f64 [N,P] and J f64 [N,N]
    (any strides).  Returns the rand() draw count."""
    sched = np.ascontiguousarray(sched, np.float64)
    perms = np.ascontiguousarray(perms, np.int64)
    assert confs.dtype == np.float64 and J.dtype == np.float64
    cs0, cs1 = _strides(confs)
    js0, js1 = _strides(J)
    d = C.c_uint64(0)
    rc = lib().orc_qa_dense(sched, len(sched), int(mcsteps), int(P), float(temp), int(N),
                            confs.ctypes.data, cs0, cs1, J.ctypes.data, js0, js1, perms,
                            rstate.st, C.byref(d))
    if rc:
        raise ValueError("orc_qa_dense rc=%d" % rc)
    return d.value


def sa_dense(sched, mcsteps, svec, J, perms, rstate):
    sched = np.ascontiguousarray(sched, np.float64)
    perms = np.ascontiguousarray(perms, np.int64)
    assert svec.dtype == np.float64 and J.dtype == np.float64
    js0, js1 = _strides(J)
    d = C.c_uint64(0)
    rc = lib().orc_sa_dense(sched, len(sched), int(mcsteps), svec.ctypes.data,
                            svec.strides[0] // 8, svec.shape[0], J.ctypes.data, js0, js1,
                            perms, rstate.st, C.byref(d))
    if rc:
        raise ValueError("orc_sa_dense rc=%d" % rc)
    return d.value


def qa_parallel(sched, mcsteps, P, temp, N, confs, nbs, nthreads):
    sched = np.ascontiguousarray(sched, np.float64)
    nbs = np.ascontiguousarray(nbs, np.float64)
    cs0, cs1 = _strides(confs)
    rc = lib().orc_qa_parallel(sched, len(sched), int(mcsteps), int(P), float(temp), int(N),
                               confs.ctypes.data, cs0, cs1, nbs, nbs.shape[1], int(nthreads))
    if rc:
        raise ValueError("orc_qa_parallel rc=%d" % rc)


def qa_fixed_seq(sched, mcsteps, P, temp, spins, nbs, perms, rstate, trace=False):
    """spins int8 [P,N] in place. Returns (draws, etrace or None)."""
    sched = np.ascontiguousarray(sched, np.float64)
    nbs = np.ascontiguousarray(nbs, np.float64)
    perms = np.ascontiguousarray(perms, np.int64)
    assert spins.dtype == np.int8 and spins.flags.c_contiguous
    et = np.zeros(len(sched)) if trace else None
    d = C.c_uint64(0)
    rc = lib().orc_qa_fixed_seq(sched, len(sched), int(mcsteps), int(P), float(temp),
                                spins.shape[1], spins, nbs, nbs.shape[1], perms, rstate.st,
                                C.byref(d), et.ctypes.data if trace else None)
    if rc:
        raise ValueError("orc_qa_fixed_seq rc=%d" % rc)
    return d.value, et


def qa_colored(sched, mcsteps, P, temp, spins, nbs, color, ncolors, seed, replica=0,
               sweep_base=0, trace=False):
    sched = np.ascontiguousarray(sched, np.float64)
    nbs = np.ascontiguousarray(nbs, np.float64)
    color = np.ascontiguousarray(color, np.int32)
    assert spins.dtype == np.int8 and spins.flags.c_contiguous
    et = np.zeros(len(sched)) if trace else None
    rc = lib().orc_qa_colored(sched, len(sched), int(mcsteps), int(P), float(temp),
                              spins.shape[1], spins, nbs, nbs.shape[1], color, int(ncolors),
                              int(seed) & 0xFFFFFFFFFFFFFFFF, int(replica), int(sweep_base),
                              et.ctypes.data if trace else None)
    if rc:
        raise ValueError("orc_qa_colored rc=%d" % rc)
    return et


def qa_colored_stats(sched, mcsteps, P, temp, spins, nbs, color, ncolors, seed, replica=0):
    """qa_colored that also returns its work: [words, stochastic words,
    stochastic attempts, Philox blocks of an early-exit comparator]."""
    sched = np.ascontiguousarray(sched, np.float64)
    nbs = np.ascontiguousarray(nbs, np.float64)
    color = np.ascontiguousarray(color, np.int32)
    assert spins.dtype == np.int8 and spins.flags.c_contiguous
    st = np.zeros(4, np.uint64)
    rc = lib().orc_qa_colored_stats(sched, len(sched), int(mcsteps), int(P), float(temp),
                                    spins.shape[1], spins, nbs, nbs.shape[1], color, int(ncolors),
                                    int(seed) & 0xFFFFFFFFFFFFFFFF, int(replica), st)
    if rc:
        raise ValueError("orc_qa_colored_stats rc=%d" % rc)
    return st


def sa_colored(sched, mcsteps, spins, nbs, color, ncolors, seed, replica_base=0, sweep_base=0):
    """Parallel-sweep classical annealing of the chains spins int8 [C, N]
    (C even) in place; see orc_sa_colored for the chain layout."""
    sched = np.ascontiguousarray(sched, np.float64)
    nbs = np.ascontiguousarray(nbs, np.float64)
    color = np.ascontiguousarray(color, np.int32)
    assert spins.dtype == np.int8 and spins.flags.c_contiguous and spins.ndim == 2
    rc = lib().orc_sa_colored(sched, len(sched), int(mcsteps), spins.shape[0], spins.shape[1],
                              spins, nbs, nbs.shape[1], color, int(ncolors),
                              int(seed) & 0xFFFFFFFFFFFFFFFF, int(replica_base), int(sweep_base))
    if rc:
        raise ValueError("orc_sa_colored rc=%d" % rc)


def sa_multispin(temps, svec, nbs, perms, U):
    """Anneal_multispin restatement on packed words svec uint64 [N] (in
    place); temps float32 [G] per sweep; perms int64 [G, N]; U float64
    [G, N, 64]."""
    temps = np.ascontiguousarray(temps, np.float32)
    nbs = np.ascontiguousarray(nbs, np.float64)
    perms = np.ascontiguousarray(perms, np.int64)
    U = np.ascontiguousarray(U, np.float64)
    assert svec.dtype == np.uint64 and svec.flags.c_contiguous
    G, N = len(temps), svec.shape[0]
    assert perms.shape == (G, N) and U.shape == (G, N, 64)
    rc = lib().orc_sa_multispin(temps, G, N, svec, nbs, nbs.shape[1], perms, U)
    if rc:
        raise ValueError("orc_sa_multispin rc=%d" % rc)


def pack_multispin(bits):
    """[64, N] 0/1 (nonzero = 1) -> uint64 [N], row k at bit 63-k
    (piqmc/sa.pyx:336-343)."""
    b = (np.asarray(bits) != 0).astype(np.uint64)
    out = np.zeros(b.shape[1], np.uint64)
    for k in range(b.shape[0]):
        out = (out << np.uint64(1)) ^ b[k]
    return out


def unpack_multispin(words):
    """uint64 [N] -> float64 [64, N] of 0/1 (piqmc/sa.pyx:402-405)."""
    w = np.asarray(words, np.uint64)
    return np.stack([((w >> np.uint64(63 - k)) & np.uint64(1)).astype(np.float64)
                     for k in range(64)])


def energy_nbs(spins, nbs):
    spins = np.ascontiguousarray(spins, np.int8)
    if spins.ndim == 1:
        spins = spins[None, :]
    nbs = np.ascontiguousarray(nbs, np.float64)
    E = np.zeros(spins.shape[0])
    lib().orc_energy_nbs(spins, spins.shape[0], spins.shape[1], nbs, nbs.shape[1], E)
    return E


def energy_coo(spins, rows, cols, vals):
    spins = np.ascontiguousarray(spins, np.float64)
    return lib().orc_energy_coo(spins, len(spins), len(vals),
                                np.ascontiguousarray(rows, np.int64),
                                np.ascontiguousarray(cols, np.int64),
                                np.ascontiguousarray(vals, np.float64))
