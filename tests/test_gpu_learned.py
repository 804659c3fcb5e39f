"""The learned proposal at the 256-spin workload: 2-D +-J L=16 (N=256), P=32,
T=1/32, 2048 replicas, MADE 256->512->512->256 trained on 131,072 slices that
plain anneals emit at Gamma <= 0.5.  Residual energy per spin against GPU wall
time for the plain sweep and for the sweep with learned moves, measured from
the committed best-known energy of the instance, and what the network gains
at equal field steps.  The measured ratio (learned time to reach the plain
method's best residual / plain time) is printed; reaching it within half the
plain time is the target, and the test records, not asserts, it.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _e_ref():
    with open(os.path.join(GOLDEN, "pmj_hit_energies.json")) as f:
        return json.load(f)["L16_seed1"]["energy"]


def test_learned_slice_moves_residual_vs_wall_time(gpu):
    from piqmc import learned
    out = learned.run(eval_steps=(25, 50, 100, 200), e_ref=_e_ref(), move="slice")
    print(json.dumps({k: out[k] for k in ("curves", "time_ratio", "train_seconds", "train_nll_per_epoch")}))
    plain, lrn = out["curves"]["plain"], out["curves"]["learned"]
    assert out["train_nll_per_epoch"][-1] < 0.5 * out["train_nll_per_epoch"][0]  # the network learned
    for p in plain + lrn:
        assert p["residual_per_spin"] >= 0.0        # nothing below the best-known energy
    assert plain[-1]["residual_per_spin"] < plain[0]["residual_per_spin"]
    # the moves are taken, and at equal field steps they leave the residual no
    # worse than the plain sweep's beyond the statistical error (3 sigma)
    assert max(q["acceptance"] for q in lrn) > 0.0
    for p, q in zip(plain, lrn):
        assert q["field_steps"] == p["field_steps"]
        tol = 3 * np.hypot(p["residual_se"], q["residual_se"])
        assert q["residual_per_spin"] <= p["residual_per_spin"] + tol, (p, q)


def test_learned_worldline_moves_after_the_anneal(gpu):
    # world-line rounds after plain anneals: accepted moves only lower the
    # energy of a slice-uniform replica, so the residual never rises
    from piqmc import learned
    out = learned.run(eval_steps=(25, 50), e_ref=_e_ref(), move="worldline", beta=2.0, rounds=(1, 2, 4))
    print(json.dumps({k: out[k] for k in ("curves", "time_ratio")}))
    plain, lrn = out["curves"]["plain"], out["curves"]["learned"]
    for p in plain:
        for q in (q for q in lrn if q["field_steps"] == p["field_steps"]):
            assert q["seconds"] > p["seconds"]
            assert q["residual_per_spin"] <= p["residual_per_spin"] + 3 * np.hypot(p["residual_se"],
                                                                                   q["residual_se"])
