"""The three larger workloads at their real size, against the oracle.

- Heterogeneous batch: the 512-instance persistent launch (2-D +-J L up to
  64 and SK N up to 256, P up to 128, S up to 1000); sample jobs (an L=64,
  P=128 lattice, the longest SK N=256 P=128 job, an SK N=128 job) checked
  replica by replica against orc_qa_colored.
- 3-D L=12 lattice with fields: the full 1024-replica, P=128, 2000-step
  launch; replicas 0 and 1023 against orc_qa_colored.
- Learned proposal network at its real shape (MADE 256->512->512->256):
  log q and gradients against the FP64 oracle, and an MH round over 2048
  replicas whose decisions agree with the action change and q ratio
  recomputed on the CPU.
The oracle runs for several replicas at once in threads (ctypes releases
the GIL).
"""
import concurrent.futures as cf

import numpy as np
import pytest

import made_oracle as MO
import oracle as O
from conftest import pm_j_problem

pytestmark = pytest.mark.gpu


def _oracle_colored(args):
    sched, P, T, sp, nbs, color, seed, replica = args
    O.qa_colored(sched, 1, P, T, sp, nbs, color, int(color.max()) + 1, seed, replica=replica)
    return sp


def test_batch_512_instances_vs_oracle(gpu):
    import torch
    from piqmc import device, workloads
    jobs = workloads.batch_mix(512, 4, 0)
    probs, b, outs = {}, device.Batch(), []
    for job in jobs:
        if job["key"] not in probs:
            n, nbs = workloads.batch_instance(job)
            probs[job["key"]] = (device.Problem(nbs), nbs)
        pb, _ = probs[job["key"]]
        c0 = workloads.batch_init(job, pb.nspins)
        c8 = c0.to(gpu)
        b.add(pb, c8, np.linspace(3.0, 1e-8, job["S"]), 1, job["T"], job["seed"],
              replica_base=job["replica_base"])
        outs.append((c0, c8))
    assert b.info()["jobs"] == 512
    b.run()
    torch.cuda.synchronize()
    # sample jobs: the shortest L=64 P=128 lattice, the longest SK N=256 P=128
    # job (982 field steps) and the shortest SK N=128 job at P=128
    pick = {}
    for job in jobs:
        for tag, ok in (("L64", job["kind"] == "pm" and job["size"] == 64 and job["P"] == 128),
                        ("SK256", job["kind"] == "sk" and job["size"] == 256 and job["P"] == 128),
                        ("SK128", job["kind"] == "sk" and job["size"] == 128 and job["P"] == 128)):
            if ok:
                best = pick.get(tag)
                longer = tag == "SK256"
                if best is None or (job["S"] > best["S"] if longer else job["S"] < best["S"]):
                    pick[tag] = job
    assert pick["SK256"]["S"] >= 950 and pick["L64"]["size"] == 64
    work, where = [], []
    for tag, job in pick.items():
        pb, nbs = probs[job["key"]]
        color = pb.coloring()
        c0, c8 = outs[job["t"]]
        for r in ((0, job["reps"] - 1) if tag != "SK256" else (0,)):
            work.append((np.linspace(3.0, 1e-8, job["S"]), job["P"], job["T"], c0[r].numpy().copy(), nbs,
                         color, job["seed"], job["replica_base"] + r))
            where.append((tag, job["t"], r, c8[r].cpu().numpy()))
    with cf.ThreadPoolExecutor(max_workers=len(work)) as ex:
        refs = list(ex.map(_oracle_colored, work))
    for (tag, t, r, got), ref in zip(where, refs):
        assert np.array_equal(got, ref), (tag, t, r)


def test_3d_field_full_launch_vs_oracle(gpu):
    # 3-D +-J L=12 (N=1728) periodic with fields h ~ N(0, 0.1^2), P=128,
    # T=1/128, 1024 replicas, 2000 field steps: the field path in one launch
    import torch
    from piqmc import device, instances
    N, r, c, v = instances.lattice_pm_j(12, 3, np.random.RandomState(5), field_sigma=0.1)
    nbs = instances.neighbors(N, r, c, v)
    pb = device.Problem(nbs)
    assert pb.pm_j_fields
    R, P, T, S = 1024, 128, 1.0 / 128, 2000
    sched = np.linspace(3.0, 1e-8, S)
    g = torch.Generator(device="cpu").manual_seed(1)
    c0 = torch.randint(0, 2, (R, P, N), generator=g, dtype=torch.int8) * 2 - 1
    c8 = c0.to(gpu)
    pb.qa_colored(c8, sched, 1, T, 3)
    torch.cuda.synchronize()
    color = pb.coloring()
    work = [(sched, P, np.float32(T), c0[k].numpy().copy(), nbs, color, 3, k) for k in (0, R - 1)]
    with cf.ThreadPoolExecutor(max_workers=2) as ex:
        refs = list(ex.map(_oracle_colored, work))
    for k, ref in zip((0, R - 1), refs):
        assert np.array_equal(c8[k].cpu().numpy(), ref), "replica %d" % k


def test_made_full_shape_vs_oracle(gpu):
    # MADE 256 -> 512 -> 512 -> 256 after a few Adam steps on +-J slices
    import torch
    from piqmc.made import MADE
    N, H = 256, 512
    m = MADE(N, H, H, seed=4)
    rng = np.random.RandomState(3)
    data = torch.tensor((2 * rng.randint(2, size=(1024, N)) - 1).astype(np.int8), device=gpu)
    data[:, :64] = 1
    for _ in range(4):
        m.train_step(data, 1e-3)
    S = (2 * rng.randint(2, size=(256, N)) - 1).astype(np.int8)
    S[:, :32] = 1
    Sd = torch.tensor(S, device=gpu)
    lq = m.logprob(Sd).cpu().numpy()
    p = m.params()
    cache, nll = MO.forward(p, S, round_bf16=True)
    np.testing.assert_allclose(lq, -nll, rtol=0, atol=2e-2 * N ** 0.5)
    m.train_step(Sd, 0.0)
    gg = m.grads()
    go = MO.backward([MO.bf16(x) if i % 2 == 0 else x for i, x in enumerate(p)], cache)
    for a, b in zip(gg, go):
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert err < 3e-2, err


def test_mh_round_2048_replicas_decisions(gpu):
    # one independence MH slice round over 2048 replicas of the 256-spin
    # lattice, P=32: every replica's slice k is replaced exactly when accepted;
    # decisions agree with exp(-dS/T) q(old)/q(new) recomputed on the CPU
    # wherever that ratio is decisively above 1 (must accept) or below 1e-30
    # (must reject)
    import torch
    from piqmc import device
    from piqmc.made import MADE, jperp
    N, nbs, _, _, _ = pm_j_problem(16, seed=1)
    pb = device.Problem(nbs)
    R, P, T, gamma = 2048, 32, 1.0 / 32, 0.3
    m = MADE(N, 512, 512, seed=6)
    g = torch.Generator(device="cpu").manual_seed(2)
    confs = (torch.randint(0, 2, (R, P, N), generator=g, dtype=torch.int8) * 2 - 1).to(gpu)
    pb.qa_colored(confs, np.linspace(3.0, gamma, 40), 1, T, 8)
    m.train_step(confs[:, 0, :].contiguous(), 1e-3)
    new, lqn = m.sample(R, 5, 0)
    k = torch.randint(0, P, (R,), generator=g, dtype=torch.int32).to(gpu)
    before = confs.clone()
    jp = jperp(gamma, P, T)
    acc = m.mh(pb, confs, new, lqn, k, jp, T, 7, 0).cpu().numpy().astype(bool)
    torch.cuda.synchronize()
    old = before[torch.arange(R), k.long()].contiguous()
    lqo = m.logprob(old).cpu().numpy()
    before, after, newc = before.cpu().numpy(), confs.cpu().numpy(), new.cpu().numpy()
    kk = k.cpu().numpy()
    # spins: slice k replaced iff accepted, everything else untouched
    for r in range(R):
        exp_r = before[r].copy()
        if acc[r]:
            exp_r[kk[r]] = newc[r]
        assert np.array_equal(after[r], exp_r), r
    # action change of the slice: classical energy plus Trotter bonds to k-1, k+1
    so, sn = before[np.arange(R), kk], newc
    nb = (before[np.arange(R), (kk - 1) % P].astype(np.float64) + before[np.arange(R), (kk + 1) % P])
    dE = O.energy_nbs(sn.copy(), nbs) - O.energy_nbs(so.copy(), nbs)
    dT = -float(np.float32(jp)) * ((sn - so.astype(np.float64)) * nb).sum(1)
    logr = -(dE + dT) / T + lqo - lqn.cpu().numpy()
    must = logr > 1e-6
    never = logr < -69
    assert np.all(acc[must]) and not np.any(acc[never])
    assert must.sum() + never.sum() > R // 2
