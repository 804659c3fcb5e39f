"""Learned slice proposal: MADE over one Trotter slice (libpiqmc_amd kernels).

``MADE`` wraps the network (BF16 MFMA forward/backward, FP32 master weights,
Adam).  ``anneal_with_proposals`` runs the parallel sweep and, every
``every`` field steps, one independence Metropolis-Hastings slice move per
replica: a slice k is drawn uniformly, a new slice x' ~ q is proposed, and it
is accepted with probability min(1, exp(-dS/T) q(x)/q(x')), where dS is the
fixed-semantics action change (classical energy of the slice plus its Trotter
bonds to k-1 and k+1).  That leaves the sweep's stationary law unchanged; the
plain sweep keeps running between moves.
"""
import ctypes as C

import numpy as np

from . import _lib
from .device import _check_dev, _stream


class MADE:
    def __init__(self, nspins, h1, h2, seed=0):
        self.nspins, self.h1, self.h2 = int(nspins), int(h1), int(h2)
        h = C.c_void_p()
        _lib.call("piqmc_made_create", self.nspins, self.h1, self.h2, int(seed), C.byref(h))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib.piqmc_made_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _shapes(self):
        N, H1, H2 = self.nspins, self.h1, self.h2
        return [(H1, N), (H1,), (H2, H1), (H2,), (N, H2), (N,)]

    def _fetch(self, fn):
        arrs = [np.zeros(s, np.float32) for s in self._shapes()]
        _lib.call(fn, self._h, *[a.ctypes.data for a in arrs])
        return arrs

    def params(self):
        """[w0, b0, w1, b1, w2, b2] as float32 numpy arrays."""
        return self._fetch("piqmc_made_get_params")

    def grads(self):
        """Gradients of the last train_step (mean NLL), same layout as params."""
        return self._fetch("piqmc_made_get_grads")

    def logprob(self, slices, stream=None):
        import torch
        B = slices.shape[0]
        _check_dev(slices, torch.int8, (B, self.nspins), "slices")
        out = torch.empty(B, dtype=torch.float64, device=slices.device)
        _lib.call("piqmc_made_logprob", self._h, B, slices.data_ptr(), out.data_ptr(), _stream(stream))
        return out

    def train_step(self, slices, lr, stream=None):
        import torch
        B = slices.shape[0]
        _check_dev(slices, torch.int8, (B, self.nspins), "slices")
        nll = C.c_double()
        _lib.call("piqmc_made_train_step", self._h, B, slices.data_ptr(), float(lr), C.byref(nll),
                  _stream(stream))
        return nll.value

    def fit(self, data, epochs=5, batch=1024, lr=1e-3, seed=0, stream=None):
        """Adam on the mean NLL of ``data`` (int8 [n, N] on the GPU), one
        native epoch call each (no host synchronisation between steps).
        Returns the per-epoch mean training NLL."""
        import torch
        g = torch.Generator(device="cpu").manual_seed(seed)
        n = data.shape[0]
        _check_dev(data, torch.int8, (n, self.nspins), "data")
        data = data.contiguous()
        hist = []
        for _ in range(epochs):
            perm = torch.randperm(n, generator=g).to(device=data.device, dtype=torch.int32)
            nll = C.c_double()
            _lib.call("piqmc_made_train_epoch", self._h, n, data.data_ptr(), perm.data_ptr(), int(batch),
                      float(lr), C.byref(nll), _stream(stream))
            hist.append(nll.value)
        return hist

    def set_graph(self, use_graph):
        """Run the sampler's N-step launch chain as a captured HIP graph
        (default) or as eager launches."""
        _lib.call("piqmc_made_set_graph", self._h, int(bool(use_graph)))

    def sample(self, B, seed, round_, device="cuda", stream=None):
        import torch
        s = torch.empty((B, self.nspins), dtype=torch.int8, device=device)
        lq = torch.empty(B, dtype=torch.float64, device=device)
        _lib.call("piqmc_made_sample", self._h, int(B), int(seed), int(round_), s.data_ptr(),
                  lq.data_ptr(), _stream(stream))
        return s, lq

    def ar_sample(self, B, seed, round_, beta=1.0, device="cuda", stream=None):
        """B draws from q_beta by the fused sampler: (int8 [B, N], log q float64 [B])."""
        import torch
        s = torch.empty((B, self.nspins), dtype=torch.int8, device=device)
        lq = torch.empty(B, dtype=torch.float64, device=device)
        _lib.call("piqmc_made_ar", self._h, int(B), 0, None, float(beta), int(seed), int(round_),
                  s.data_ptr(), lq.data_ptr(), _stream(stream))
        return s, lq

    def ar_logprob(self, slices, beta=1.0, stream=None):
        """log q_beta of the given slices, by the fused sampler's arithmetic."""
        import torch
        B = slices.shape[0]
        _check_dev(slices, torch.int8, (B, self.nspins), "slices")
        lq = torch.empty(B, dtype=torch.float64, device=slices.device)
        _lib.call("piqmc_made_ar", self._h, 0, int(B), slices.data_ptr(), float(beta), 0, 0, None,
                  lq.data_ptr(), _stream(stream))
        return lq

    def worldline_mh(self, problem, confs, beta, temp, seed, round_, stream=None):
        """One world-line move per replica of confs int8 [R,P,N] (in place):
        a slice y ~ q_beta replaces slice 0, the same spins flip in every
        slice.  Returns the int32 [R] accepted flags."""
        import torch
        R, P, N = confs.shape
        _check_dev(confs, torch.int8, (R, P, self.nspins), "confs")
        acc = torch.zeros(R, dtype=torch.int32, device=confs.device)
        _lib.call("piqmc_made_worldline_mh", self._h, problem._h, R, P, confs.data_ptr(), float(beta),
                  float(temp), int(seed), int(round_), acc.data_ptr(), _stream(stream))
        return acc

    def slice_mh(self, problem, confs, slice_idx, beta, jperp, temp, seed, round_, stream=None):
        """The independence slice move with a draw of the fused sampler: slice
        slice_idx[r] of replica r is replaced by y ~ q_beta, accepted with
        min(1, exp(-dS/T) q_beta(old)/q_beta(y)).  Returns int32 [R] flags."""
        import torch
        R, P, N = confs.shape
        _check_dev(confs, torch.int8, (R, P, self.nspins), "confs")
        _check_dev(slice_idx, torch.int32, (R,), "slice_idx")
        acc = torch.zeros(R, dtype=torch.int32, device=confs.device)
        _lib.call("piqmc_made_slice_mh", self._h, problem._h, R, P, confs.data_ptr(), slice_idx.data_ptr(),
                  float(beta), float(jperp), float(temp), int(seed), int(round_), acc.data_ptr(),
                  _stream(stream))
        return acc

    def mh(self, problem, confs, new, logq_new, slice_idx, jperp, temp, seed, round_, stream=None):
        """One MH slice move per replica of confs int8 [R,P,N] (in place).
        Returns the int32 [R] accepted flags."""
        import torch
        R, P, N = confs.shape
        _check_dev(confs, torch.int8, (R, P, self.nspins), "confs")
        _check_dev(new, torch.int8, (R, self.nspins), "new")
        _check_dev(slice_idx, torch.int32, (R,), "slice_idx")
        acc = torch.zeros(R, dtype=torch.int32, device=confs.device)
        _lib.call("piqmc_made_mh", self._h, problem._h, R, P, confs.data_ptr(), new.data_ptr(),
                  logq_new.data_ptr(), slice_idx.data_ptr(), float(jperp), float(temp), int(seed),
                  int(round_), acc.data_ptr(), _stream(stream))
        return acc


def jperp(gamma, P, temp):
    """J_perp of one schedule entry (piqmc/qmc.pyx:95), float32."""
    pt = np.float32(P) * np.float32(temp)
    return float(np.float32(-0.5 * P * float(np.float32(temp)) * np.log(np.tanh(gamma / float(pt)))))


def anneal_with_proposals(problem, made, confs, sched, mcsteps, temp, seed, every=4,
                          stats=None, fused=True, beta=1.0):
    """Parallel sweep of confs int8 [R,P,N] over ``sched`` with a MADE slice
    move for every replica after each block of ``every`` field steps.  The
    proposal comes from the fused sampler (q_beta), or with fused=False from
    the forward-pass sampler (q)."""
    import torch
    R, P, N = confs.shape
    sched = np.asarray(sched, np.float64)
    S = len(sched)
    g = torch.Generator(device="cpu").manual_seed(int(seed) & 0x7FFFFFFF)
    rnd = 0
    for s0 in range(0, S, every):
        sub = sched[s0:s0 + every]
        problem.qa_colored(confs, sub, mcsteps, temp, seed, sweep_base=s0 * mcsteps)
        k = torch.randint(0, P, (R,), generator=g, dtype=torch.int32).to(confs.device)
        if fused:
            acc = made.slice_mh(problem, confs, k, beta, jperp(sub[-1], P, temp), temp, seed, rnd)
        else:
            new, lq = made.sample(R, seed, rnd, device=confs.device)
            acc = made.mh(problem, confs, new, lq, k, jperp(sub[-1], P, temp), temp, seed, rnd)
        if stats is not None:
            stats.append(int(acc.sum()))
        rnd += 1
    return confs
