"""Tile autotuner for the parallel sweep.

The sweep's tiles -- threads per workgroup and replicas per workgroup (each
workgroup keeps its replicas' bit-packed slices in LDS) -- change only speed:
every choice takes identical decisions (tests/test_gpu_colored.py pins this).
``tune_colored`` times each candidate that fits in LDS on a short schedule that
samples the real one (the cost per attempt depends on the field: how many bits
need a random draw), keeps the fastest, and caches it in a JSON file keyed by
device architecture and problem shape.
"""
import json
import os

import numpy as np

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "piqmc", "tune.json")
BLOCKS = (64, 128, 256, 512, 1024)
RPWS = (1, 2, 4, 8, 16)
LDS_BYTES = 160 * 1024


def _cache_path():
    return os.environ.get("PIQMC_TUNE_CACHE", DEFAULT_CACHE)


def _load():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(db):
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def _arch():
    import torch
    p = torch.cuda.get_device_properties(0)
    return getattr(p, "gcnArchName", p.name)


def key(pb, P, R):
    path = 1 if pb.pm_j else (2 if pb.pm_j_fields else 0)
    return "%s|N%d|Z%d|zmax%d|pmj%d|nc%d|P%d|R%d" % (_arch(), pb.nspins, pb.maxnb, pb.zmax,
                                                     path, pb.ncolors, P, R)


def candidates(pb, P, R):
    word = 4 if P // 2 <= 32 else 8
    out = []
    for rpw in RPWS:
        if rpw > R or 2 * pb.nspins * word * rpw + 1200 > LDS_BYTES:
            continue
        for block in BLOCKS:
            if block // 64 > max(1, pb.nspins // 64) * 2:
                continue           # more waves than words to update
            out.append((block, rpw))
    return out or [(64, 1)]


def tune_colored(pb, P, R, temp, sched, mcsteps=1, samples=8, use_cache=True, verbose=False):
    """Best (block, rpw) for sweeping R replicas of P slices of problem pb."""
    import torch
    from . import device
    k = key(pb, P, R)
    db = _load() if use_cache else {}
    if k in db:
        return tuple(db[k]["tile"])
    sched = np.asarray(sched, np.float64)
    sub = sched[np.linspace(0, len(sched) - 1, min(samples, len(sched))).astype(int)]
    g = torch.Generator(device="cpu").manual_seed(1)
    confs0 = (torch.randint(0, 2, (R, P, pb.nspins), generator=g, dtype=torch.int8) * 2 - 1).cuda()
    confs = torch.empty_like(confs0)
    timing = {}
    for block, rpw in candidates(pb, P, R):
        best = float("inf")
        for _ in range(2):
            confs.copy_(confs0)
            pb.qa_colored(confs, sub, mcsteps, temp, 7, block=block, rpw=rpw)
            best = min(best, device.last_kernel_ms())
        timing["%d,%d" % (block, rpw)] = best
        if verbose:
            print("tile block=%d rpw=%d: %.3f ms" % (block, rpw, best))
    tile = min(timing, key=timing.get)
    block, rpw = (int(x) for x in tile.split(","))
    db[k] = {"tile": [block, rpw], "ms": timing}
    if use_cache:
        _store(db)
    return block, rpw
