"""Algorithmic VALU work of the parallel +-1 sweep on the 1024-spin workload
(test infrastructure: run by hand, writes profiles/colored_pmj1024_work.md).

    python oracle/work_model.py [--replicas K]

The oracle's colored sweep (orc_qa_colored_stats) counts, for sampled replicas
of the benchmark's own inputs (the seeded +-J instance, the torch-seeded
initial spins and Philox seed 12345), the parity words swept, the words that
hold a stochastic attempt, and the Philox4x32-10 blocks that an MSB-first
comparator with per-word early exit consumes.  The minimal VALU work per
word follows from the bit-sliced algorithm itself (one 32-bit lane op = one
VALU lane-instruction), counted in the summary this script writes:

  deterministic part, every word (z = 4 neighbours, 32 slices per word):
    4  bond masks s ^ n_j ^ neg_j          (one three-input XOR each)
    6  4-input bit-sliced adder            (full adder = XOR3 + MAJ3, then a
                                            half adder, then 2 carry ops)
    1  rotation of the other parity word   (Trotter neighbours k-1 / k+1)
    2  Trotter bond masks                  (s ^ other, s ^ rotated other)
    2  Trotter count bits                  (XOR, AND)
    3  "count >= k" masks for k = 1..3     (k = 4 is the top bit)
    2  accept mask                         (two 2:1 selects by the Trotter count)
    2  stochastic mask                     (likewise)
    1  flip                                (s ^ accept)
   = 23
  per Philox block (four 32-bit uniform bit planes):
   40  ten rounds x (two 32x32->64 products + two three-input XORs); the key
       schedule is wave-uniform (scalar unit)
    8  comparator: per plane, update the accepted and undecided masks
   = 48
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(ROOT, "pathintegral-qmc_amd"))

DET_OPS_PER_WORD = 23
OPS_PER_BLOCK = 48


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--replicas", type=int, default=8)
    args = ap.parse_args()
    import torch
    import oracle as O
    from piqmc import instances   # pure-python instance generator (no device)

    L, P, T, S, R = 32, 64, np.float32(1.0 / 64), 500, 4096
    N, r, c, v = instances.lattice_pm_j(L, 2, np.random.RandomState(2024))
    nbs = instances.neighbors(N, r, c, v)
    # the problem handle's 2-coloring: breadth-first from spin 0 (color 0)
    color = np.array([((i // L) + (i % L)) % 2 for i in range(N)], np.int32)
    g = torch.Generator(device="cpu").manual_seed(0)
    init = (torch.randint(0, 2, (R, P, N), generator=g, dtype=torch.int8) * 2 - 1).numpy()
    sched = np.linspace(3.0, 1e-8, S)
    reps = np.linspace(0, R - 1, args.replicas).astype(int)
    tot = np.zeros(4, np.float64)
    t0 = time.time()
    for rep in reps:
        sp = np.ascontiguousarray(init[rep])
        tot += O.qa_colored_stats(sched, 1, P, T, sp, nbs, color, 2, 12345, replica=int(rep))
    dt = time.time() - t0
    words, stoch_words, stoch_att, blocks = tot / len(reps)
    assert words == 2 * N * S
    words_launch = words * R
    ops_word = DET_OPS_PER_WORD + OPS_PER_BLOCK * blocks / words
    key = "R=%d N=%d P=%d S=%d" % (R, N, P, S)
    out = os.path.join(ROOT, "profiles", "colored_pmj1024_work.md")
    with open(out, "w") as f:
        f.write("# Algorithmic VALU work of the parallel sweep: 1024-spin +-J workload, 4096 replicas, "
                "500 field steps\n\n")
        f.write("Command (run from the repository root, CPU only, %.0f s):\n\n"
                "    python oracle/work_model.py --replicas %d\n\n" % (dt, args.replicas))
        f.write("Replicas sampled (the benchmark's own initial spins and Philox seed): %s\n\n"
                % ", ".join(str(x) for x in reps))
        f.write("| per replica (mean) | value |\n|---|---|\n")
        f.write("| parity words swept (32 attempts each) | %d |\n" % words)
        f.write("| words holding a stochastic attempt | %.0f (%.2f %%) |\n" % (stoch_words, 100 * stoch_words / words))
        f.write("| stochastic attempts | %.0f (%.3f per word) |\n" % (stoch_att, stoch_att / words))
        f.write("| Philox blocks, per-word early exit | %.0f (%.4f per word) |\n\n" % (blocks, blocks / words))
        f.write("Minimal VALU lane-operations per word = %d (deterministic) + %d x %.4f (Philox blocks) = %.2f.\n"
                % (DET_OPS_PER_WORD, OPS_PER_BLOCK, blocks / words, ops_word))
        f.write("Per launch: %.4g words x %.2f = %.4g lane-operations = %.4g wave64 instructions.\n\n"
                % (words_launch, ops_word, words_launch * ops_word, words_launch * ops_word / 64))
        f.write("words_per_launch: %.0f %s\n" % (words_launch, key))
        f.write("min_valu_lane_ops_per_word: %.4f %s\n" % (ops_word, key))
    print(open(out).read())


if __name__ == "__main__":
    main()
