"""Preanneal -> tile -> quantum anneal -> lowest-energy slice, on the device.

examples/spinglass32.py:94-137 runs sa.Anneal on one spin vector, tiles the
vector over the Trotter slices, runs qmc.QuantumAnneal on the slices and
keeps the slice of lowest classical energy.  :func:`anneal` runs that
pipeline for R replicas at once, every stage on the GPU with the state
resident in HBM between stages:

1. random +-1 spins [R, N] (or the caller's);
2. classical preanneal of the R spin vectors with the parallel sweep
   (``Problem.sa_colored``: zero Trotter coupling, temperature per step);
3. tile each vector over P slices ([R, P, N]);
4. quantum anneal with the parallel sweep (``Problem.qa_colored``);
5. classical energy of every slice, and per replica the slice of lowest
   energy.

(The example builds its slice tiling before the preanneal, so its quantum
anneal starts from the random vector; here the tiled state is the
preannealed one, which is what the example's comments describe.)

The preanneal draws from Philox key ``2*seed + 1`` and the quantum anneal
from key ``2*seed``, so the two stages use independent streams.
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class PipelineResult:
    confs: object          # int8 [R, P, N] final slices (device)
    energies: object       # float64 [R, P] classical energy of every slice (device)
    best_energy: object    # float64 [R] lowest slice energy per replica (device)
    best_conf: object      # int8 [R, N] the slice that has it (device)


def random_spins(replicas, nspins, seed, device):
    """Seeded +-1 int8 [replicas, nspins] (drawn on the host, then copied)."""
    import torch
    g = torch.Generator(device="cpu").manual_seed(int(seed))
    s = torch.randint(0, 2, (replicas, nspins), generator=g, dtype=torch.int8) * 2 - 1
    return s.to(device)


def anneal(problem, replicas, slices, temp, field_sched, pre_sched=None, mcsteps=1,
           pre_mcsteps=1, seed=0, init=None, block=0, rpw=0, stream=None):
    """Run the pipeline for ``replicas`` replicas of ``slices`` slices.

    problem     -- piqmc.device.Problem
    temp        -- temperature of the quantum anneal
    field_sched -- transverse fields of the quantum anneal
    pre_sched   -- temperatures of the classical preanneal (None: skip)
    init        -- optional int8 [replicas, nspins] device tensor of +-1
                   starting spins (consumed in place)
    """
    import torch
    R, P, N = int(replicas), int(slices), problem.nspins
    if R < 1:
        raise ValueError("replicas must be >= 1")
    dev = torch.device("cuda", torch.cuda.current_device())
    seed = int(seed) & 0x7FFFFFFFFFFFFFFF
    if init is None:
        spins = random_spins(R, N, seed, dev)
    else:
        if tuple(init.shape) != (R, N) or init.dtype != torch.int8 or not init.is_cuda:
            raise ValueError("init must be an int8 CUDA tensor of shape (%d, %d)" % (R, N))
        spins = init.contiguous()
    if pre_sched is not None and len(pre_sched) > 0:
        # sa_colored pairs chains; an odd count gets one scratch chain
        chains = spins if R % 2 == 0 else torch.cat([spins, spins[:1]], 0)
        problem.sa_colored(chains, pre_sched, pre_mcsteps, 2 * seed + 1, block=block, rpw=rpw,
                           stream=stream)
        spins = chains[:R]
    confs = spins[:, None, :].expand(R, P, N).contiguous()
    problem.qa_colored(confs, field_sched, mcsteps, temp, 2 * seed, block=block, rpw=rpw,
                       stream=stream)
    E = problem.slice_energies(confs, stream=stream)
    best, idx = E.min(1)
    best_conf = confs[torch.arange(R, device=dev), idx]
    return PipelineResult(confs=confs, energies=E, best_energy=best, best_conf=best_conf)
