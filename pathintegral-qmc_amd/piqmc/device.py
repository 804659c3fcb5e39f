"""Batched device API: many replicas resident in HBM, one launch per call.

A :class:`Problem` holds a neighbour table (tools.GenerateNeighbors format) in
HBM together with its graph coloring.  Its methods take torch tensors that
already live on the GPU (torch is used only for device memory and streams):

* ``qa_replay``   -- qmc.QuantumAnneal for R replicas, bit-exact replay
                     (piqmc/qmc.pyx:30-136);
* ``sa_replay``   -- sa.Anneal for R replicas (piqmc/sa.pyx:50-120);
* ``qa_colored``  -- the graph-colored, parity-alternated parallel sweep;
* ``sa_colored``  -- the same sweep as classical annealing (zero Trotter
                     coupling, a temperature per step);
* ``slice_energies`` -- classical energy of every slice.

Spin tensors are int8 +1/-1: replicas x slices x spins ([R, P, N]).
"""
import ctypes as C

import numpy as np

from . import _lib


def _stream(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _check_dev(t, dtype, shape, name):
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("%s must be a CUDA (HIP) torch tensor" % name)
    if t.dtype != dtype:
        raise ValueError("%s must have dtype %s, got %s" % (name, dtype, t.dtype))
    if tuple(t.shape) != tuple(shape):
        raise ValueError("%s must have shape %s, got %s" % (name, tuple(shape), tuple(t.shape)))
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % name)


def last_kernel_ms():
    """Device milliseconds of the main sweep kernel of the last replay /
    parallel call from this thread (HIP events on that call's stream)."""
    ms = C.c_float()
    _lib.call("piqmc_last_kernel_ms", C.byref(ms))
    return ms.value


def rand_seed(seed):
    """glibc srand(seed) image of the TYPE_3 rand() state, uint32[32]."""
    st = np.zeros(32, np.uint32)
    _lib.call("piqmc_rand_seed", int(seed) & 0xFFFFFFFF, st.ctypes.data)
    return st


def rand_states(seeds):
    return np.stack([rand_seed(s) for s in seeds])


def perm_streams(seeds, nspins, nsweeps, nthreads=0):
    """uint16 [nsweeps, nspins, R]: stream r is what QuantumAnneal draws from
    numpy RandomState(seeds[r]) (first permutation of range(N), then chained)."""
    seeds = np.ascontiguousarray(seeds, np.uint32)
    out = np.empty((nsweeps, nspins, len(seeds)), np.uint16)
    _lib.call("piqmc_perm_streams", seeds.ctypes.data, len(seeds), int(nspins), int(nsweeps),
              out.ctypes.data, int(nthreads))
    return out


class Problem:
    """A neighbour table resident in HBM (piqmc_problem handle)."""

    def __init__(self, nbs):
        nbs = np.ascontiguousarray(nbs, np.float64)
        if nbs.ndim != 3 or nbs.shape[2] != 2:
            raise ValueError("nbs must be [nspins, maxnb, 2]")
        self.nspins, self.maxnb = nbs.shape[0], nbs.shape[1]
        h = C.c_void_p()
        _lib.call("piqmc_problem_create", self.nspins, self.maxnb, nbs.ctypes.data, C.byref(h))
        self._h = h
        self._load_info()

    @classmethod
    def from_dense(cls, J):
        """A problem from a dense coupling matrix (float64 [N, N]; diagonal =
        fields; the upper triangle is read), the argument of
        QuantumAnneal_dense / Anneal_dense.  Replays on it follow those
        functions (double-precision couplings; sa replay accepts ediff > 0)."""
        J = np.asarray(J)
        if J.dtype != np.float64 or J.ndim != 2 or J.shape[0] != J.shape[1]:
            raise ValueError("J must be a square float64 matrix")
        self = cls.__new__(cls)
        self.nspins = J.shape[0]
        h = C.c_void_p()
        s0, s1 = (s // 8 for s in J.strides)
        _lib.call("piqmc_problem_create_dense", self.nspins, J.ctypes.data, s0, s1, C.byref(h))
        self._h = h
        self._load_info()
        return self

    def _load_info(self):
        vals = [C.c_int() for _ in range(5)]
        _lib.call("piqmc_problem_info", self._h, *[C.byref(v) for v in vals])
        self.maxnb = vals[1].value
        self.ncolors, self.zmax = vals[2].value, vals[4].value
        self.pm_j = vals[3].value == 1            # +-1 couplings only
        self.pm_j_fields = vals[3].value == 2     # +-1 couplings and fields
        ex = C.c_int()
        _lib.call("piqmc_problem_exact_rows", self._h, C.byref(ex))
        self.exact_rows = bool(ex.value)          # local fields sum exactly in any order

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib.piqmc_problem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def coloring(self):
        """The spatial coloring in use, int32 [nspins]."""
        out = np.zeros(self.nspins, np.int32)
        _lib.call("piqmc_problem_get_coloring", self._h, out.ctypes.data)
        return out

    def set_sweep_path(self, path):
        """Parallel-sweep kernel: 'auto', 'word' (a parity word per thread) or
        'dense' (a replica per wavefront, a lane per slice)."""
        code = {"auto": 0, "word": 1, "dense": 2}[path]
        _lib.call("piqmc_problem_set_sweep_path", self._h, code)

    DENSE_VARIANTS = ("auto", "wave", "split", "workgroup", "pipe4", "pipe8", "fields")

    def set_dense_variant(self, variant):
        """Layout of the dense sweep: 'auto', 'wave' (a replica per wavefront,
        a lane per slice), 'split' (a slice's field over several lanes),
        'workgroup' (a replica per workgroup), 'pipe4' / 'pipe8' (that
        workgroup pipelined with three / seven summing wavefronts), 'fields'
        (every local field cached in LDS and updated on flips)."""
        code = self.DENSE_VARIANTS.index(variant)
        _lib.call("piqmc_problem_set_dense_variant", self._h, code)
        self.dense_variant = variant

    def set_coloring(self, color):
        color = np.ascontiguousarray(color, np.int32)
        nc = int(color.max()) + 1
        _lib.call("piqmc_problem_set_coloring", self._h, nc, color.ctypes.data)
        self.ncolors = nc

    # ------------------------------------------------------------ replay
    def qa_replay(self, confs, sched, mcsteps, temp, perms, rand_state, stream=None):
        """confs int8 [R,P,N] in place; perms int16/uint16 [S*M, N, R];
        rand_state int32 [R, 32] in place.  Returns (draws int64 [R], ties int32 [R])."""
        import torch
        R, P, N = confs.shape
        sched = np.ascontiguousarray(sched, np.float64)
        S, M = len(sched), int(mcsteps)
        _check_dev(confs, torch.int8, (R, P, self.nspins), "confs")
        if perms.dtype not in (torch.int16, torch.uint16):
            raise ValueError("perms must be 16-bit")
        _check_dev(perms, perms.dtype, (S * M, N, R), "perms")
        _check_dev(rand_state, torch.int32, (R, 32), "rand_state")
        draws = torch.zeros(R, dtype=torch.int64, device=confs.device)
        ties = torch.zeros(R, dtype=torch.int32, device=confs.device)
        _lib.call("piqmc_qa_replay", self._h, R, P, S, M, float(temp), sched.ctypes.data,
                  confs.data_ptr(), perms.data_ptr(), rand_state.data_ptr(), draws.data_ptr(),
                  ties.data_ptr(), _stream(stream))
        return draws, ties

    def sa_replay(self, svec, sched, mcsteps, perms, rand_state, stream=None):
        """svec int8 [R,N] in place; perms [S*M, N, R]; rand_state int32 [R,32]."""
        import torch
        R, N = svec.shape
        sched = np.ascontiguousarray(sched, np.float64)
        S, M = len(sched), int(mcsteps)
        _check_dev(svec, torch.int8, (R, self.nspins), "svec")
        _check_dev(perms, perms.dtype, (S * M, N, R), "perms")
        _check_dev(rand_state, torch.int32, (R, 32), "rand_state")
        draws = torch.zeros(R, dtype=torch.int64, device=svec.device)
        ties = torch.zeros(R, dtype=torch.int32, device=svec.device)
        _lib.call("piqmc_sa_replay", self._h, R, S, M, sched.ctypes.data, svec.data_ptr(),
                  perms.data_ptr(), rand_state.data_ptr(), draws.data_ptr(), ties.data_ptr(),
                  _stream(stream))
        return draws, ties

    # ------------------------------------------------------------ parallel
    def qa_colored(self, confs, sched, mcsteps, temp, seed, replica_base=0, sweep_base=0,
                   trace=False, block=0, rpw=0, stream=None):
        """Parallel sweep of confs int8 [R,P,N] in place.  Returns the
        per-field-step mean slice energy float64 [R,S] when trace=True."""
        import torch
        R, P, N = confs.shape
        sched = np.ascontiguousarray(sched, np.float64)
        S = len(sched)
        _check_dev(confs, torch.int8, (R, P, self.nspins), "confs")
        et = torch.zeros((R, S), dtype=torch.float64, device=confs.device) if trace else None
        _lib.call("piqmc_qa_colored", self._h, R, P, S, int(mcsteps), float(temp),
                  sched.ctypes.data, confs.data_ptr(), int(seed) & 0xFFFFFFFFFFFFFFFF,
                  int(replica_base), int(sweep_base), et.data_ptr() if trace else None,
                  int(block), int(rpw), _stream(stream))
        return et

    def sa_colored(self, spins, sched, mcsteps, seed, replica_base=0, sweep_base=0, block=0,
                   rpw=0, stream=None):
        """Parallel-sweep classical annealing of the chains spins int8 [C, N]
        (C even) in place, temperature sched[s] in step s."""
        import torch
        C, N = spins.shape
        sched = np.ascontiguousarray(sched, np.float64)
        _check_dev(spins, torch.int8, (C, self.nspins), "spins")
        _lib.call("piqmc_sa_colored", self._h, C, len(sched), int(mcsteps), sched.ctypes.data,
                  spins.data_ptr(), int(seed) & 0xFFFFFFFFFFFFFFFF, int(replica_base),
                  int(sweep_base), int(block), int(rpw), _stream(stream))

    # ------------------------------------------------------------ energies
    def slice_energies(self, spins, stream=None):
        """Classical energy of each row of int8 spins [..., N] (float64)."""
        import torch
        flat = spins.reshape(-1, self.nspins)
        _check_dev(flat, torch.int8, flat.shape, "spins")
        E = torch.zeros(flat.shape[0], dtype=torch.float64, device=spins.device)
        _lib.call("piqmc_slice_energies", self._h, flat.shape[0], flat.data_ptr(), E.data_ptr(),
                  _stream(stream))
        return E.reshape(spins.shape[:-1])


class Batch:
    """Heterogeneous parallel sweeps in one persistent launch.

    ``add`` takes the arguments of :meth:`Problem.qa_colored`; ``run`` sweeps
    every job with one launch whose workgroups take work items from a device
    queue.  Each job's spins (and energy trace) equal its own ``qa_colored``
    call with the same seed, replica base and sweep base.
    """

    def __init__(self):
        h = C.c_void_p()
        _lib.call("piqmc_batch_create", C.byref(h))
        self._h = h
        self._keep = []
        self.jobs = []

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                _lib.lib.piqmc_batch_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def add(self, problem, confs, sched, mcsteps, temp, seed, replica_base=0, sweep_base=0,
            trace=False):
        import torch
        R, P, N = confs.shape
        sched = np.ascontiguousarray(sched, np.float64)
        _check_dev(confs, torch.int8, (R, P, problem.nspins), "confs")
        et = torch.zeros((R, len(sched)), dtype=torch.float64, device=confs.device) if trace else None
        _lib.call("piqmc_batch_add", self._h, problem._h, R, P, len(sched), int(mcsteps), float(temp),
                  sched.ctypes.data, confs.data_ptr(), int(seed) & 0xFFFFFFFFFFFFFFFF,
                  int(replica_base), int(sweep_base), et.data_ptr() if trace else None)
        self._keep.append((problem, confs, et))
        self.jobs.append({"problem": problem, "R": R, "P": P, "sched": sched, "mcsteps": int(mcsteps),
                          "temp": float(temp), "seed": int(seed), "replica_base": int(replica_base),
                          "sweep_base": int(sweep_base)})
        return et

    def run(self, block=0, rpw=0, stream=None, tune=False):
        """One launch over every job.  With ``tune`` and rpw == 0, the
        replicas per work item come from the autotuner (piqmc.autotune.tune_batch)."""
        if tune and not rpw:
            from . import autotune
            rpw = autotune.tune_batch(self)
        _lib.call("piqmc_batch_run", self._h, int(block), int(rpw), _stream(stream))

    def prepare(self, block=0, rpw=0, stream=None):
        """Upload the tables and build the work queue (then ``launch``)."""
        _lib.call("piqmc_batch_prepare", self._h, int(block), int(rpw), _stream(stream))

    def launch(self, stream=None):
        """Pack, sweep every job, unpack: kernel launches only."""
        _lib.call("piqmc_batch_launch", self._h, _stream(stream))

    def capture(self, block=0, rpw=0):
        """A torch CUDAGraph (a HIP graph) of ``launch``: each replay runs every
        job's schedule again from the spins the previous run left."""
        import torch
        self.prepare(block, rpw)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        with torch.cuda.graph(g, stream=side):
            self.launch(stream=side)
        return g

    def info(self):
        v = [C.c_int() for _ in range(3)]
        _lib.call("piqmc_batch_info", self._h, *[C.byref(x) for x in v])
        return {"jobs": v[0].value, "items": v[1].value, "grid": v[2].value}
