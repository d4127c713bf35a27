"""Tuning sweeps and planner evaluation.

    python -m convgemm_amd.tuning --train 2000 --test 256 --dtype bf16 \
        --db convgemm_amd/data/mi355x_bf16_fwd.db --model convgemm_amd/data/mi355x_bf16_fwd.mlp

1. draws random forward problems from the batched-workload ranges
   (batch {1,4,16}, H=W {7,14,28,56}, Cin/Cout 32..1024 step 32, kernel
   {1,3,5}, stride {1,2}, pad k//2, NHWC);
2. autotunes every training problem (exhaustive over tile configuration x
   split-K, all timings kept in the tuning database);
3. trains the MLP cost model on those timings with the device kernels;
4. autotunes the held-out test problems too (the exhaustive reference) and
   reports, per planner, the regret t(chosen plan) / t(best plan) measured
   from the recorded timings: analytic ("no learning") vs learned model.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

import convgemm_amd as cg

DTYPES = {"f32": cg.F32, "bf16": cg.BF16, "fp8": cg.FP8_E4M3}


def random_problems(n, seed, batches=(1, 4, 16), hws=(7, 14, 28, 56), chans=range(32, 1025, 32), ks=(1, 3, 5),
                    strides=(1, 2)):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        k = int(rng.choice(ks))
        out.append(dict(b=int(rng.choice(batches)), h=int(rng.choice(hws)), c=int(rng.choice(list(chans))),
                        kn=int(rng.choice(list(chans))), k=k, s=int(rng.choice(strides)), pad=k // 2))
    return out


def desc_of(p, layout=cg.NHWC, op=cg.FWD):
    return cg.make_desc(op, layout, p["b"], p["h"], p["h"], p["c"], p["kn"], p["k"], p["k"], p["pad"], p["pad"],
                        p["s"], p["s"])


class Buffers:
    """Device operands reused across problems (grown on demand)."""

    def __init__(self, torch):
        self.torch = torch
        self.x = self.w = self.y = None
        self.dtype = None

    def get(self, p, dtype):
        t = self.torch
        ho = (p["h"] + 2 * p["pad"] - p["k"]) // p["s"] + 1
        nx, nw, ny = p["b"] * p["h"] * p["h"] * p["c"], p["c"] * p["k"] * p["k"] * p["kn"], p["b"] * ho * ho * p["kn"]
        if self.x is None or self.x.numel() < nx or self.dtype != dtype:
            self.x = self._rand(max(nx, 1 << 20), dtype)
        if self.w is None or self.w.numel() < nw or self.dtype != dtype:
            self.w = self._rand(max(nw, 1 << 20), dtype)
        if self.y is None or self.y.numel() < ny:
            self.y = t.zeros(max(ny, 1 << 20), device="cuda")
        self.dtype = dtype
        return self.w, self.x, self.y

    def _rand(self, n, dtype):
        """Random operands (timings on all-zero data run at unrepresentative clocks)."""
        t = self.torch
        r = t.randn(n, device="cuda")
        if dtype == cg.BF16:
            return r.to(t.bfloat16).view(t.int16)
        if dtype == cg.FP8_E4M3:
            return r.to(t.float8_e4m3fn).view(t.uint8)
        return r


def tune_all(problems, dtype, reps=3, log_every=100):
    import torch
    bufs = Buffers(torch)
    scales = (1.0, 1.0) if dtype == cg.FP8_E4M3 else None
    t0 = time.time()
    for i, p in enumerate(problems):
        w, x, y = bufs.get(p, dtype)
        cg.autotune(desc_of(p), w, x, y, dtype=dtype, scales=scales, reps=reps)
        if log_every and (i + 1) % log_every == 0:
            print(f"tuned {i + 1}/{len(problems)} in {time.time() - t0:.1f} s", flush=True)


def regret(problems, dtype, planner):
    """Mean and max of t(plan chosen by `planner`) / t(best), from recorded timings."""
    cg.set_planner(planner)
    rs = []
    for p in problems:
        d = desc_of(p)
        rec = {(c, s): us for c, s, us in cg.tune_get(d, dtype)}
        if not rec:
            continue
        c, s = cg.plan(d, dtype)
        best = min(rec.values())
        t = rec.get((c, s))
        if t is None:  # plan outside the tuned candidate set: charge the slowest recorded time
            t = max(rec.values())
        rs.append(t / best)
    cg.set_planner(cg.PLAN_TUNED)
    rs = np.array(rs)
    return {"mean": float(rs.mean()), "geomean": float(np.exp(np.log(rs).mean())), "max": float(rs.max()),
            "p90": float(np.percentile(rs, 90)), "n": int(rs.size)}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--test", type=int, default=256)
    ap.add_argument("--dtype", default="bf16", choices=list(DTYPES))
    ap.add_argument("--epochs", type=int, default=2000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--db", default="")
    ap.add_argument("--model", default="")
    a = ap.parse_args(argv)
    dtype = DTYPES[a.dtype]
    cg.lib()
    cg.tune_clear()
    train = random_problems(a.train, a.seed)
    test = random_problems(a.test, a.seed + 1)
    t0 = time.time()
    tune_all(train, dtype)
    t_tune = time.time() - t0
    t0 = time.time()
    loss = cg.costmodel_train(epochs=a.epochs, lr=a.lr, seed=a.seed)
    t_train = time.time() - t0
    if a.model:
        cg.costmodel_save(a.model)
    tune_all(test, dtype, log_every=0)  # exhaustive reference for the held-out problems
    res = {
        "dtype": a.dtype, "train_problems": a.train, "test_problems": a.test, "problems_in_db": cg.tune_count(),
        "tune_seconds": round(t_tune, 1), "train_seconds": round(t_train, 2), "final_train_loss": loss,
        "regret_test_analytic": regret(test, dtype, cg.PLAN_ANALYTIC),
        "regret_test_model": regret(test, dtype, cg.PLAN_MODEL),
        "regret_train_model": regret(train[:256], dtype, cg.PLAN_MODEL),
    }
    if a.db:
        cg.tune_save(a.db)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
