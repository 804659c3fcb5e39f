"""Plain vs learned-proposal annealing at equal GPU wall time.

Default workload: 2-D +-J glass L=16 (N=256), P=32, T=1/32 (P*T=1), 2048 replicas,
Gamma linear 3.0 -> 1e-8.  A MADE 256->512->512->256 (BF16, FP32 accumulation,
Adam lr 1e-3) is trained on ~1e5 slices that plain anneals emit at Gamma <= 0.5.
Two learned methods use it as a Metropolis-Hastings proposal, drawn by the
fused sampler from q_beta (beta = 1: the network itself):
  - "slice" (default): the independence slice move every ``every`` field
    steps during the anneal (the Trotter bonds of the replaced slice enter
    the acceptance);
  - "worldline": after a plain anneal of S field steps, rounds of the
    world-line move (a drawn slice replaces slice 0 and the same spins flip in
    every slice, so no Trotter bond changes); timed as the plain anneal plus
    the rounds.
Residual energy per spin = mean over replicas of (min over slices E - E_ref) / N,
with E_ref the committed best-known energy of the instance when given, else the
lowest energy any run reached.  ``dim`` and ``field_sigma`` select other
lattices (the 3-D L=12 field workload uses dim=3, field_sigma=0.1, P=128, 1024
replicas and a 1728->1024->1024->1728 network).
"""
import time

import numpy as np


def _timed(fn):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return out, time.perf_counter() - t0


def run(L=16, P=32, R=2048, hidden=512, train_steps=500, epochs=5, lr=1e-3,
        eval_steps=(25, 50, 100, 200), every=4, seed=0, device="cuda", dim=2,
        field_sigma=0.0, move="slice", beta=1.0, rounds=(1, 2, 4, 8), e_ref=None):
    import torch
    from . import device as D, instances
    from .made import MADE, anneal_with_proposals

    N, r, c, v = instances.lattice_pm_j(L, dim, np.random.RandomState(seed + 1), field_sigma)
    pb = D.Problem(instances.neighbors(N, r, c, v))
    T = 1.0 / P
    g = torch.Generator(device="cpu").manual_seed(seed)

    def rand_confs():
        return (torch.randint(0, 2, (R, P, N), generator=g, dtype=torch.int8) * 2 - 1).to(device)

    # 1. slices emitted by a plain anneal at Gamma <= 0.5
    sched = np.linspace(3.0, 1e-8, train_steps)
    cut = int(np.searchsorted(-sched, -0.5))
    confs = rand_confs()
    (_, t_data) = _timed(lambda: pb.qa_colored(confs, sched[:cut], 1, T, seed + 11))
    snaps = []
    pos = cut
    for stop in (cut + (train_steps - cut) // 2, train_steps):
        (_, dt) = _timed(lambda: pb.qa_colored(confs, sched[pos:stop], 1, T, seed + 11, sweep_base=pos))
        t_data += dt
        pos = stop
        snaps.append(confs.reshape(-1, N).clone())
    data = torch.cat(snaps)
    # 2. training
    made = MADE(N, hidden, hidden, seed=seed + 5)
    hist, t_train = _timed(lambda: made.fit(data, epochs=epochs, batch=1024, lr=lr, seed=seed))
    # 3. evaluation at matched schedules
    best = np.inf if e_ref is None else float(e_ref)
    plain, learned = [], []
    for S in eval_steps:
        sch = np.linspace(3.0, 1e-8, S)
        init = rand_confs()
        cp = init.clone()
        _, tp = _timed(lambda: pb.qa_colored(cp, sch, 1, T, seed + 100 + S))
        ep = pb.slice_energies(cp).min(1).values.cpu().numpy()
        # replicas whose slices all agree (no kink in imaginary time)
        kinkfree = float((cp == cp[:, :1]).all(dim=2).all(dim=1).float().mean())
        plain.append((S, tp, ep, kinkfree))
        if move == "worldline":
            cl = cp.clone()
            tl, accs, k = tp, 0, 0
            for kmax in rounds:
                while k < kmax:
                    acc, dt = _timed(lambda: made.worldline_mh(pb, cl, beta, T, seed + 7 * S, k))
                    tl += dt
                    accs += int(acc.sum())
                    k += 1
                el = pb.slice_energies(cl).min(1).values.cpu().numpy()
                learned.append((S, k, tl, el, accs / (k * R)))
                best = min(best, el.min()) if e_ref is None else best
        else:
            cl = init.clone()
            stats = []
            _, tl = _timed(lambda: anneal_with_proposals(pb, made, cl, sch, 1, T, seed + 100 + S,
                                                         every=every, stats=stats, beta=beta))
            el = pb.slice_energies(cl).min(1).values.cpu().numpy()
            learned.append((S, len(stats), tl, el, float(np.sum(stats)) / max(1, len(stats) * R)))
            best = min(best, el.min()) if e_ref is None else best
        if e_ref is None:
            best = min(best, ep.min())
    def se(e):
        return float(np.std(e, ddof=1) / np.sqrt(len(e)) / N)
    res = {"plain": [{"field_steps": S, "seconds": tp, "residual_per_spin": float((ep - best).mean() / N),
                      "residual_se": se(ep), "kink_free_fraction": kf}
                     for S, tp, ep, kf in plain],
           "learned": [{"field_steps": S, "rounds": k, "seconds": tl,
                        "residual_per_spin": float((el - best).mean() / N), "residual_se": se(el),
                        "acceptance": acc}
                       for S, k, tl, el, acc in learned]}
    # time for the learned method to reach the plain method's best residual
    target = min(p["residual_per_spin"] for p in res["plain"])
    t_plain = min(p["seconds"] for p in res["plain"] if p["residual_per_spin"] <= target)
    hits = [q["seconds"] for q in res["learned"] if q["residual_per_spin"] <= target]
    out = {"workload": "%d-D +-J L=%d N=%d%s, P=%d, T=1/%d, %d replicas, MADE %d->%d->%d->%d" % (
        dim, L, N, ", h~N(0,%g^2)" % field_sigma if field_sigma else "", P, P, R, N, hidden, hidden, N),
        "move": move, "beta": beta,
        "train_slices": int(data.shape[0]), "train_nll_per_epoch": hist, "train_seconds": t_train,
        "train_data_seconds": t_data,
        # forward 2(N H1 + H1 H2 + H2 N) per slice, backward about twice that
        "train_tflops": 3 * 2.0 * (N * hidden + hidden * hidden + hidden * N) * int(data.shape[0]) * epochs
                        / t_train / 1e12,
        "e_ref_per_spin": float(best / N), "e_ref": "committed best-known energy" if e_ref is not None
        else "lowest energy reached by any run", "curves": res,
        "plain_time_to_best": t_plain,
        "learned_time_to_plain_best": min(hits) if hits else None,
        "time_ratio": (min(hits) / t_plain) if hits else None}
    return out


if __name__ == "__main__":
    import json
    print(json.dumps(run()))
