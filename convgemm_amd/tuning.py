"""Tuning sweeps and planner evaluation.

    CONVGEMM_COSTMODEL=none python -m convgemm_amd.tuning --train 150 --test 32 \
        --mix f32:fwd:nhwc,f32:fwd:nchw,f32:wgrad:nhwc,f32:dgrad:nhwc,bf16:fwd:nhwc:2000,bf16:fwd:nchw,bf16:wgrad:nhwc,bf16:dgrad:nhwc,fp8:fwd:nhwc \
        --db mix.db --model convgemm_amd/data/mi355x_costmodel.mlp

is the command behind the shipped model (CONVGEMM_COSTMODEL=none: the run
starts with no model loaded).

1. draws random problems from the batched-workload ranges (batch {1,4,16},
   H=W {7,14,28,56}, Cin/Cout 32..1024 step 32, kernel {1,3,5}, stride
   {1,2}, pad k//2) for each (operand type, operation, layout) of --mix;
2. autotunes every training problem (exhaustive over tile configuration x
   split-K, all timings kept in the tuning database);
3. trains the MLP cost model on those timings with the device kernels;
4. autotunes the held-out test problems too (the exhaustive reference) and
   reports, per planner and per --mix entry, the regret t(chosen plan) /
   t(best plan) measured from the recorded timings: analytic ("no
   learning") vs learned model.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

import convgemm_amd as cg

DTYPES = {"f32": cg.F32, "bf16": cg.BF16, "fp8": cg.FP8_E4M3}
OPS = {"fwd": cg.FWD, "wgrad": cg.WGRAD, "dgrad": cg.DGRAD}
LAYOUTS = {"nhwc": cg.NHWC, "nchw": cg.NCHW}


def random_problems(n, seed, batches=(1, 4, 16), hws=(7, 14, 28, 56), chans=range(32, 1025, 32), ks=(1, 3, 5),
                    strides=(1, 2), dtype=cg.BF16, op=cg.FWD, layout=cg.NHWC):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        k = int(rng.choice(ks))
        out.append(dict(b=int(rng.choice(batches)), h=int(rng.choice(hws)), c=int(rng.choice(list(chans))),
                        kn=int(rng.choice(list(chans))), k=k, s=int(rng.choice(strides)), pad=k // 2,
                        dtype=dtype, op=op, layout=layout))
    return out


def desc_of(p, layout=None, op=None):
    layout = p.get("layout", cg.NHWC) if layout is None else layout
    op = p.get("op", cg.FWD) if op is None else op
    return cg.make_desc(op, layout, p["b"], p["h"], p["h"], p["c"], p["kn"], p["k"], p["k"], p["pad"], p["pad"],
                        p["s"], p["s"])


class Buffers:
    """Device operands reused across problems (grown on demand)."""

    def __init__(self, torch):
        self.torch = torch
        self.x = self.w = self.y = None
        self.dtype = None

    def get(self, p, dtype):
        """(in, x, out) operands for p's operation: (W, X, Y), (dY, X, dW) or (W, dY, dX)."""
        t = self.torch
        ho = (p["h"] + 2 * p["pad"] - p["k"]) // p["s"] + 1
        nx, nw, ny = p["b"] * p["h"] * p["h"] * p["c"], p["c"] * p["k"] * p["k"] * p["kn"], p["b"] * ho * ho * p["kn"]
        n = max(nx, nw, ny)
        if self.x is None or self.x.numel() < n or self.dtype != dtype:
            self.x = self._rand(max(n, 1 << 20), dtype)
        if self.w is None or self.w.numel() < n or self.dtype != dtype:
            self.w = self._rand(max(n, 1 << 20), dtype)
        if self.y is None or self.y.numel() < n:
            self.y = t.zeros(max(n, 1 << 20), device="cuda")
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


def tune_all(problems, dtype=None, reps=3, log_every=100):
    """Autotune every problem (each carries its own operand type unless `dtype` is given)."""
    import torch
    bufs = Buffers(torch)
    t0 = time.time()
    for i, p in enumerate(problems):
        dt = p.get("dtype", cg.BF16) if dtype is None else dtype
        scales = (1.0, 1.0) if dt == cg.FP8_E4M3 else None
        w, x, y = bufs.get(p, dt)
        cg.autotune(desc_of(p), w, x, y, dtype=dt, scales=scales, reps=reps)
        if log_every and (i + 1) % log_every == 0:
            print(f"tuned {i + 1}/{len(problems)} in {time.time() - t0:.1f} s", flush=True)


def regret(problems, dtype, planner):
    """Mean and max of t(plan chosen by `planner`) / t(best), from recorded timings
    (dtype None: each problem's own operand type)."""
    cg.set_planner(planner)
    rs, outside = [], 0
    for p in problems:
        dt = p.get("dtype", cg.BF16) if dtype is None else dtype
        d = desc_of(p)
        rec = {(c, s): us for c, s, us in cg.tune_get(d, dt)}
        if not rec:
            continue
        c, s = cg.plan(d, dt)
        best = min(rec.values())
        t = rec.get((c, s))
        if t is None:  # plan outside the tuned candidate set: charge the slowest recorded time
            t = max(rec.values())
            outside += 1
        rs.append(t / best)
    cg.set_planner(cg.PLAN_TUNED)
    rs = np.array(rs)
    return {"mean": float(rs.mean()), "geomean": float(np.exp(np.log(rs).mean())), "max": float(rs.max()),
            "p90": float(np.percentile(rs, 90)), "n": int(rs.size), "outside": outside}


def parse_mix(spec, default_train=None):
    """'f32:fwd:nhwc,bf16:fwd:nhwc:2000' -> [(name, dtype, op, layout, n_train)]: an optional fourth
    field gives that entry's training-problem count (default `default_train`)."""
    out = []
    for item in spec.split(","):
        f = item.split(":")
        if len(f) not in (3, 4):
            raise ValueError(f"mix entry {item!r}: expected dtype:op:layout[:n_train]")
        n = int(f[3]) if len(f) == 4 else default_train
        out.append((":".join(f[:3]), DTYPES[f[0]], OPS[f[1]], LAYOUTS[f[2]], n))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=2000,
                    help="training problems per --mix entry without its own count")
    ap.add_argument("--test", type=int, default=256, help="held-out problems per --mix entry")
    ap.add_argument("--mix", default="bf16:fwd:nhwc", help="comma-separated dtype:op:layout[:n_train] entries")
    ap.add_argument("--epochs", type=int, default=2000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--db", default="")
    ap.add_argument("--model", default="")
    a = ap.parse_args(argv)
    mix = parse_mix(a.mix, a.train)
    cg.lib()
    cg.tune_clear()
    train, test = [], {}
    for i, (name, dt, op, lay, n_train) in enumerate(mix):
        train += random_problems(n_train, a.seed + 10 * i, dtype=dt, op=op, layout=lay)
        test[name] = random_problems(a.test, a.seed + 10 * i + 1, dtype=dt, op=op, layout=lay)
    t0 = time.time()
    tune_all(train)
    t_tune = time.time() - t0
    t0 = time.time()
    loss = cg.costmodel_train(epochs=a.epochs, lr=a.lr, seed=a.seed, every=100,
                              progress=lambda ep, l: print(f"train epoch {ep}/{a.epochs} loss {l:.4f} "
                                                           f"({time.time() - t0:.1f} s)", flush=True))
    t_train = time.time() - t0
    if a.model:
        cg.costmodel_save(a.model)
    for name in test:
        tune_all(test[name], log_every=100)  # exhaustive reference for the held-out problems
    res = {"mix": a.mix, "train_problems": len(train), "test_problems": sum(len(v) for v in test.values()),
           "problems_in_db": cg.tune_count(), "tune_seconds": round(t_tune, 1), "train_seconds": round(t_train, 2),
           "final_train_loss": loss, "per_mix": {}}
    for name, probs in test.items():
        res["per_mix"][name] = {"analytic": regret(probs, None, cg.PLAN_ANALYTIC),
                                "model": regret(probs, None, cg.PLAN_MODEL)}
    all_test = [p for v in test.values() for p in v]
    res["regret_test_analytic"] = regret(all_test, None, cg.PLAN_ANALYTIC)
    res["regret_test_model"] = regret(all_test, None, cg.PLAN_MODEL)
    res["regret_train_model"] = regret(train[::max(1, len(train) // 256)], None, cg.PLAN_MODEL)
    if a.db:
        cg.tune_save(a.db)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
