"""Learned cost model: device MLP kernels, training, persistence, planning.

* Gradients of the device kernels (forward, MSE, backward, weight
  reductions) against the FP64 oracle in oracle/mlp.py, which is itself
  pinned by finite differences in the CPU suite: rel <= 1e-4 of max|g|.
* Training on autotuner timings lowers the standardized loss, the model
  survives a save/load round trip bit-for-bit, and the planner built on it
  runs convolutions that match the tuned plan's results.
* On held-out problems, the plans the model chooses are measured against the
  exhaustive autotuner's best (regret = t(plan) / t(best)) and compared with
  the analytic planner, the no-learning method.
"""
import json
import os

import numpy as np
import pytest
import torch

from oracle import mlp as omlp

pytestmark = pytest.mark.gpu

import convgemm_amd as cg  # noqa: E402
from convgemm_amd import tuning  # noqa: E402
from tests.gpu_util import DEV  # noqa: E402


SHIPPED_MODEL = os.path.join(os.path.dirname(cg.__file__), "data", "mi355x_costmodel.mlp")


@pytest.fixture(autouse=True)
def _clean_tuning_state():
    """Each test starts from an empty database; afterwards the shipped model and the
    start-up planner are back, for the tests that follow."""
    start = cg.get_planner()
    cg.tune_clear()
    cg.set_planner(cg.PLAN_TUNED)
    yield
    cg.tune_clear()
    cg.costmodel_load(SHIPPED_MODEL)
    cg.set_planner(start)


@pytest.mark.parametrize("n", [1, 37, 300])
def test_device_gradients_match_oracle(n):
    rng = np.random.default_rng(100 + n)
    p = (rng.standard_normal(omlp.NP) * 0.2).astype(np.float32)
    X = rng.standard_normal((n, omlp.F)).astype(np.float32)
    T = rng.standard_normal(n).astype(np.float32)
    loss, g = cg.mlp_loss_grad(p, X, T)
    L_ref, g_ref = omlp.loss_grad(p, X, T)
    assert abs(loss - L_ref) <= 1e-5 * max(1.0, abs(L_ref)), (loss, L_ref)
    err = np.max(np.abs(g.astype(np.float64) - g_ref)) / max(np.max(np.abs(g_ref)), 1e-30)
    assert err <= 1e-4, err
    # deterministic reductions: the same inputs give the same bits
    loss2, g2 = cg.mlp_loss_grad(p, X, T)
    assert loss2 == loss and np.array_equal(g, g2)


def test_train_requires_timings():
    with pytest.raises(cg.ConvGemmError):
        cg.costmodel_train(epochs=10)


def _tuned(problems, dtype=cg.BF16):
    tuning.tune_all(problems, dtype, reps=3, log_every=0)
    return problems


SMALL = dict(batches=(1, 4), hws=(7, 14, 28), chans=range(32, 513, 32))


def test_train_save_load_and_plan(tmp_path):
    probs = _tuned(tuning.random_problems(24, seed=5, **SMALL))
    n_rows = sum(len(cg.tune_get(tuning.desc_of(p), cg.BF16)) for p in probs)
    assert n_rows >= 24
    short = cg.costmodel_train(epochs=5, lr=3e-3, seed=1)
    seen = []
    loss = cg.costmodel_train(epochs=600, lr=3e-3, seed=1, every=200, progress=lambda ep, l: seen.append((ep, l)))
    assert [ep for ep, _ in seen] == [200, 400, 600] and all(np.isfinite(l) for _, l in seen), seen
    assert seen[-1][1] < seen[0][1] < short, (seen, short)
    assert np.isfinite(loss) and loss < short and loss < 0.5, (short, loss)  # targets are standardized: 1.0 = mean
    d = tuning.desc_of(probs[0])
    pred = cg.costmodel_predict(d, cg.BF16)
    assert len(pred) == len(cg.tune_get(d, cg.BF16)) and all(u > 0 for _, _, u in pred)
    path = str(tmp_path / "model.mlp")
    cg.costmodel_save(path)
    cg.costmodel_train(epochs=3, lr=3e-3, seed=9)  # overwrite the live model
    cg.costmodel_load(path)
    assert cg.costmodel_predict(d, cg.BF16) == pred
    with pytest.raises(cg.ConvGemmError):
        cg.costmodel_load(str(tmp_path / "missing.mlp"))

    # the MODEL planner runs what it plans, and the result matches the tuned plan's
    p = dict(b=2, h=14, c=64, kn=96, k=3, s=1, pad=1)
    dsc = tuning.desc_of(p)
    cg.set_planner(cg.PLAN_MODEL)
    c_m, s_m = cg.plan(dsc, cg.F32)
    rng = np.random.default_rng(3)
    x = torch.from_numpy(rng.standard_normal((2, 14, 14, 64)).astype(np.float32)).to(DEV)
    w = torch.from_numpy(rng.standard_normal((64 * 9, 96)).astype(np.float32)).to(DEV)
    y_m = torch.empty(2 * 14 * 14 * 96, device=DEV)
    y_a = torch.empty_like(y_m)
    cg.conv_ex(dsc, w, x, y_m)
    torch.cuda.synchronize()
    assert cg.last_plan()[:2] == (c_m, s_m)
    cg.set_planner(cg.PLAN_ANALYTIC)
    cg.conv_ex(dsc, w, x, y_a)
    torch.cuda.synchronize()
    ref = y_a.cpu().double().numpy()
    assert np.max(np.abs(y_m.cpu().double().numpy() - ref)) <= 1e-5 * np.max(np.abs(ref))


def test_learned_planner_regret_vs_analytic():
    """Held-out regret of model-chosen plans vs analytic plans (BF16 forward, NHWC)."""
    train = _tuned(tuning.random_problems(160, seed=11, **SMALL))
    test = _tuned(tuning.random_problems(48, seed=12, **SMALL))
    cg.costmodel_train(epochs=1500, lr=3e-3, seed=0)
    r_model = tuning.regret(test, cg.BF16, cg.PLAN_MODEL)
    r_an = tuning.regret(test, cg.BF16, cg.PLAN_ANALYTIC)
    r_fit = tuning.regret(train[:48], cg.BF16, cg.PLAN_MODEL)
    out = os.environ.get("CONVGEMM_RESULTS_DIR")  # where to keep the regret figures, if anywhere
    if out:
        os.makedirs(out, exist_ok=True)
        with open(os.path.join(out, "costmodel_regret.json"), "w") as f:
            json.dump({"model": r_model, "analytic": r_an, "model_train": r_fit}, f)
    print("regret model", r_model, "analytic", r_an, "fit", r_fit, flush=True)
    assert r_model["n"] == 48 and r_an["n"] == 48
    assert r_fit["geomean"] < 1.15, r_fit
    assert r_model["geomean"] < r_an["geomean"], (r_model, r_an)


SHIPPED_KINDS = ["f32:fwd:nchw", "f32:dgrad:nhwc", "bf16:wgrad:nhwc", "fp8:fwd:nhwc"]


def test_shipped_model_beats_analytic():
    """The model shipped with the library (the start-up default) on 256 fresh shapes
    from its training ranges (64 of each of four kinds, from a seed it was not
    trained on): every plan it picks is among the timed candidates, and on EVERY
    kind its geomean regret against exhaustive tuning is below the analytic
    ("no learning") planner's."""
    probs = {}
    for mix in SHIPPED_KINDS:
        (_, dt, op, lay, _), = tuning.parse_mix(mix, 0)
        probs[mix] = tuning.random_problems(64, seed=4242, dtype=dt, op=op, layout=lay)
    pooled = [p for v in probs.values() for p in v]
    tuning.tune_all(pooled, reps=3, log_every=32)
    cg.costmodel_load(SHIPPED_MODEL)
    r_m = tuning.regret(pooled, None, cg.PLAN_MODEL)
    r_a = tuning.regret(pooled, None, cg.PLAN_ANALYTIC)
    print("pooled regret model", r_m, "analytic", r_a, flush=True)
    assert r_m["n"] == len(pooled) and r_m["outside"] == 0
    worse = []
    for mix, ps in probs.items():
        m = tuning.regret(ps, None, cg.PLAN_MODEL)
        a = tuning.regret(ps, None, cg.PLAN_ANALYTIC)
        print(mix, "regret model", m["geomean"], "analytic", a["geomean"], flush=True)
        if not m["geomean"] < a["geomean"]:
            worse.append((mix, m["geomean"], a["geomean"]))
    assert not worse, worse
