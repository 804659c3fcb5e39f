"""Tile autotuner for the parallel sweep.

The sweep's tiles change only speed: every choice takes identical decisions
(the GPU tests pin this).  They are
  - on the word-per-thread path: threads per workgroup and replicas per
    workgroup (each workgroup keeps its replicas' bit-packed slices in LDS;
    a lane owns one parity word, i.e. all P/2 slices of one parity of one
    spin, so the slices per lane are fixed by the layout);
  - on the dense (lane-per-slice) path: the layout of a replica's slices over
    lanes and wavefronts (a lane per slice in one wavefront, a slice's field
    over several lanes, a workgroup per replica, pipelined or not) and the
    threads per workgroup;
  - in a persistent batch: the replicas per work item of its word-per-thread
    jobs.
``tune_colored`` times each candidate on a short schedule that samples the
real one (the cost per attempt depends on the field: how many bits need a
random draw), keeps the fastest, and caches it in a JSON file keyed by device
architecture and problem shape; ``tune_batch`` does the same for a batch on
schedules shortened in proportion.
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
    return "%s|N%d|Z%d|zmax%d|pmj%d|nc%d|ex%d|P%d|R%d" % (_arch(), pb.nspins, pb.maxnb, pb.zmax,
                                                          path, pb.ncolors, int(pb.exact_rows), P, R)


def is_dense(pb):
    """The automatic path choice of the library: dense when color classes
    average fewer than 8 spins."""
    return pb.nspins < 8 * pb.ncolors


def candidates(pb, P, R):
    """(block, rpw) tiles of the word path, or (block, variant) of the dense path."""
    if is_dense(pb):
        out = [(b, "wave") for b in (64, 128, 256) if b // 64 <= max(1, R)]
        if pb.exact_rows:
            W = P // 2
            if W <= 16 and 16 % W == 0:
                out += [(b, "split") for b in (64, 128, 256) if b // 64 <= max(1, R)]
            out += [(256, "workgroup"), (256, "pipe4"), (512, "pipe8"), (64, "fields"), (256, "fields")]
        return out
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
    """Best tile for sweeping R replicas of P slices of problem pb: (block,
    rpw) on the word path.  On the dense path it returns (block, 0) and leaves
    the winning layout set on pb (Problem.set_dense_variant)."""
    import torch
    from . import device
    k = key(pb, P, R)
    db = _load() if use_cache else {}
    dense = is_dense(pb)
    if k in db:
        tile = db[k]["tile"]
        if dense:
            pb.set_dense_variant(tile[1])
            return int(tile[0]), 0
        return tuple(tile)
    sched = np.asarray(sched, np.float64)
    sub = sched[np.linspace(0, len(sched) - 1, min(samples, len(sched))).astype(int)]
    g = torch.Generator(device="cpu").manual_seed(1)
    confs0 = (torch.randint(0, 2, (R, P, pb.nspins), generator=g, dtype=torch.int8) * 2 - 1).cuda()
    confs = torch.empty_like(confs0)
    timing = {}
    for block, second in candidates(pb, P, R):
        best = float("inf")
        try:
            if dense:
                pb.set_dense_variant(second)
            for _ in range(2):
                confs.copy_(confs0)
                pb.qa_colored(confs, sub, mcsteps, temp, 7, block=block, rpw=0 if dense else second)
                best = min(best, device.last_kernel_ms())
        except ValueError:
            continue                      # the variant does not apply to this shape
        timing["%d,%s" % (block, second)] = best
        if verbose:
            print("tile block=%d %s: %.3f ms" % (block, second, best))
    tile = min(timing, key=timing.get)
    b, second = tile.split(",")
    block = int(b)
    if dense:
        pb.set_dense_variant(second)
        db[k] = {"tile": [block, second], "ms": timing}
        out = (block, 0)
    else:
        db[k] = {"tile": [block, int(second)], "ms": timing}
        out = (block, int(second))
    if use_cache:
        _store(db)
    return out


def batch_key(batch):
    shapes = sorted("%d/%d/%d/%d/%d/%d/%d" % (j["problem"].nspins, j["problem"].maxnb,
                                             int(j["problem"].pm_j), j["problem"].ncolors, j["P"], j["R"],
                                             len(j["sched"])) for j in batch.jobs)
    import hashlib
    return "%s|batch|%s" % (_arch(), hashlib.sha1("|".join(shapes).encode()).hexdigest()[:16])


def tune_batch(batch, rpws=(1, 2, 4), shorten=32, use_cache=True, verbose=False):
    """Best replicas per work item for the word-per-thread jobs of a batch
    (device.Batch): every candidate runs a copy of the batch on fresh spins
    with each schedule shortened ``shorten`` times (at least two field steps,
    sampled over the real one), so the jobs keep their relative lengths."""
    import torch
    from . import device
    k = batch_key(batch)
    db = _load() if use_cache else {}
    if k in db:
        return int(db[k]["rpw"])
    g = torch.Generator(device="cpu").manual_seed(2)
    trial = device.Batch()
    for j in batch.jobs:
        S = len(j["sched"])
        n = max(2, S // shorten)
        sub = j["sched"][np.linspace(0, S - 1, n).astype(int)] if S else j["sched"]
        confs = (torch.randint(0, 2, (j["R"], j["P"], j["problem"].nspins), generator=g,
                               dtype=torch.int8) * 2 - 1).cuda()
        trial.add(j["problem"], confs, sub, j["mcsteps"], j["temp"], j["seed"], j["replica_base"],
                  j["sweep_base"])
    timing = {}
    for rpw in rpws:
        best = float("inf")
        for _ in range(2):
            trial.run(rpw=rpw)
            best = min(best, device.last_kernel_ms())
        timing[str(rpw)] = best
        if verbose:
            print("batch rpw=%d: %.3f ms" % (rpw, best))
    rpw = int(min(timing, key=timing.get))
    db[k] = {"rpw": rpw, "ms": timing}
    if use_cache:
        _store(db)
    return rpw
