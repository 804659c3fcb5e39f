"""Synthetic workloads shared by the benchmark and the tests.

``batch_mix`` is the heterogeneous batch: 512 instances mixing 2-D +-J
lattices L in {8, 16, 32, 64} and Sherrington-Kirkpatrick glasses N in {64,
128, 256} with J_ij ~ N(0, 1/N), P in {16, 32, 64, 128}, schedule lengths
uniform in [100, 1000], T = 1/P, a few distinct instances per size.
"""
import numpy as np

from . import instances

BATCH_KINDS = (("pm", 8), ("pm", 16), ("pm", 32), ("pm", 64), ("sk", 64), ("sk", 128), ("sk", 256))


def batch_mix(n_inst=512, reps=4, seed=0):
    """Job specs: dicts with t (job index), kind ("pm" lattice or "sk"),
    size (L or N), P, S (field steps), T, iseed (instance seed), key (the
    instance's identity: jobs with equal keys share one instance), seed
    (Philox key), replica_base, reps and init_seed (torch seed of the
    initial spins)."""
    rng = np.random.RandomState(seed)
    jobs = []
    for t in range(n_inst):
        kind, size = BATCH_KINDS[rng.randint(len(BATCH_KINDS))]
        P = int(rng.choice([16, 32, 64, 128]))
        S = int(rng.randint(100, 1001))
        jobs.append({"t": t, "kind": kind, "size": size, "P": P, "S": S, "T": 1.0 / P,
                     "key": (kind, size, t % 4),
                     "iseed": (1000 if kind == "sk" else 0) + 10 * size + t % 4,
                     "seed": 1000 + t, "replica_base": t * reps, "reps": reps, "init_seed": t})
    return jobs


def batch_instance(job):
    """(N, neighbour table) of a batch job's instance."""
    if job["kind"] == "sk":
        n, r, c, v = instances.sk(job["size"], np.random.RandomState(job["iseed"]))
    else:
        n, r, c, v = instances.lattice_pm_j(job["size"], 2, np.random.RandomState(job["iseed"]))
    return n, instances.neighbors(n, r, c, v)


def batch_init(job, nspins):
    """The job's initial spins, int8 [reps, P, N] on the CPU."""
    import torch
    g = torch.Generator(device="cpu").manual_seed(job["init_seed"])
    return torch.randint(0, 2, (job["reps"], job["P"], nspins), generator=g, dtype=torch.int8) * 2 - 1
