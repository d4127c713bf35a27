#!/usr/bin/env python3
"""Benchmark: FP32 implicit-GEMM convolution on one MI355X.

Workload (one "step"): the ResNet-50-shaped layer sweep at batch 32 --
L1 1x1 256->64 @56^2, L2 3x3 64->64 @56^2, L3 3x3 s2 128->128 56->28,
L4 1x1 1024->256 @14^2, L5 3x3 512->512 @7^2 -- forward in both NHWC and NCHW
with the fused bias + batch-norm + ReLU epilogue: 10 convolutions, 57.6 GFLOP
(2*P*K*N, padding zeros counted as in plots/convgemm.plt:56-57).
Inputs are synthetic (X ~ N(0,1), He-scaled weights, epilogue vectors as in
the workload description) and resident in HBM before timing starts.

Plans: by default the tuning databases shipped in convgemm_amd/data (the
autotuner's timings of these ten problems, of their backward, and of the
batch-64 precision sweep, on an MI355X: the plans the committed profiles were
measured with) are loaded read-only; problems they lack, or every problem
with --fresh-tune, are autotuned at setup (untimed).

Prints one JSON line:  value = TFLOP/s over the timed steps, the roofline
of the dominant kernel measured with HIP events on the launch stream, the
FP64 error of every layer, and a CPU baseline (the oracle's FP32
restatement of the reference algorithm, timed on the host cores).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
SHIPPED_TUNE_DB = os.path.join(ROOT, "convgemm_amd", "data", "mi355x_fp32_sweep_b32.tunedb")
# backward (weight gradient, backward data; FP32 and BF16) of the same layers at batch 32
SHIPPED_BWD_DB = os.path.join(ROOT, "convgemm_amd", "data", "mi355x_backward_b32.tunedb")
# forward NHWC of the same layers at batch 64 in FP32, BF16 and FP8 (the precision sweep)
SHIPPED_PREC_DB = os.path.join(ROOT, "convgemm_amd", "data", "mi355x_precision_b64.tunedb")
sys.path.insert(0, ROOT)

METRIC = "conv TFLOP/s (% of roofline), BF16/FP8 vs FP32 plain at stated FP64 error"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
PEAK_TFLOPS = {"fp32": PEAK_FP32_TFLOPS, "bf16": 2500.0, "fp8": 5000.0}  # MI355X dense MFMA
ESIZE = {"fp32": 4, "bf16": 2, "fp8": 1}


def roofline_of(flops, nbytes, dname, ms):
    """Attainable TFLOP/s = min(dense MFMA peak of the operand type, AI x 8 TB/s) for
    `nbytes` compulsory HBM bytes, and the fractions of it and of the peak reached in `ms`."""
    tf = flops / (ms * 1e-3) / 1e12
    hbm = flops / nbytes * PEAK_HBM_GBS * 1e9 / 1e12
    att = min(PEAK_TFLOPS[dname], hbm)
    return {"bound": "mfma" if att == PEAK_TFLOPS[dname] else "hbm", "attainable_tflops": round(att, 1),
            "frac_roofline": round(tf / att, 3), "frac_peak": round(tf / PEAK_TFLOPS[dname], 4)}


def op_bytes(spec, op, dname):
    """Compulsory HBM bytes of one operation with operands of type dname and an FP32 output:
    forward |X| + |W| (+ |Y|), weight gradient |X| + |dY| (+ |dW|), backward data |dY| + |W|
    (+ 2|dX|: read and written, it accumulates)."""
    b, h, w, c, kn, kh, kw, pad, st = spec
    ho, wo = out_hw(h, pad, kh, st), out_hw(w, pad, kw, st)
    nx, nw, ny = b * h * w * c, c * kh * kw * kn, b * ho * wo * kn
    e = ESIZE[dname]
    if op == "fwd":
        return e * (nx + nw) + 4.0 * ny
    if op == "wgrad":
        return e * (nx + ny) + 4.0 * nw
    return e * (ny + nw) + 8.0 * nx

# name: (b, h, w, c, kn, kh, kw, pad, stride)
LAYERS = [
    ("L1_1x1_256to64_56", (32, 56, 56, 256, 64, 1, 1, 0, 1)),
    ("L2_3x3_64to64_56", (32, 56, 56, 64, 64, 3, 3, 1, 1)),
    ("L3_3x3s2_128to128_56", (32, 56, 56, 128, 128, 3, 3, 1, 2)),
    ("L4_1x1_1024to256_14", (32, 14, 14, 1024, 256, 1, 1, 0, 1)),
    ("L5_3x3_512to512_7", (32, 7, 7, 512, 512, 3, 3, 1, 1)),
]


def out_hw(h, pad, k, stride):
    return (h + 2 * pad - (k - 1) - 1) // stride + 1


def layer_flops(spec):
    b, h, w, c, kn, kh, kw, pad, st = spec
    ho, wo = out_hw(h, pad, kh, st), out_hw(w, pad, kw, st)
    return 2.0 * b * ho * wo * c * kh * kw * kn


def layer_bytes(spec):
    """Compulsory HBM bytes of a forward conv: X + W + Y + 5 epilogue vectors (FP32)."""
    b, h, w, c, kn, kh, kw, pad, st = spec
    ho, wo = out_hw(h, pad, kh, st), out_hw(w, pad, kw, st)
    return 4.0 * (b * h * w * c + c * kh * kw * kn + b * ho * wo * kn + 5 * kn)


def make_inputs(torch, layout, spec, seed):
    b, h, w, c, kn, kh, kw, pad, st = spec
    g = torch.Generator(device="cuda").manual_seed(seed)
    ho, wo = out_hw(h, pad, kh, st), out_hw(w, pad, kw, st)
    xs = (b, h, w, c) if layout == 0 else (b, c, h, w)
    ws = (c, kh, kw, kn) if layout == 0 else (kn, c, kh, kw)
    ys = (b, ho, wo, kn) if layout == 0 else (b, kn, ho, wo)
    d = dict(
        X=torch.randn(xs, device="cuda", generator=g),
        W=torch.randn(ws, device="cuda", generator=g) * (2.0 / (c * kh * kw)) ** 0.5,
        Y=torch.empty(ys, device="cuda"),
        bias=0.1 * torch.randn(kn, device="cuda", generator=g),
        mean=0.1 * torch.randn(kn, device="cuda", generator=g),
        inv_std=1.0 / torch.sqrt(0.5 + 1.5 * torch.rand(kn, device="cuda", generator=g) + 1e-5),
        gamma=0.5 + torch.rand(kn, device="cuda", generator=g),
        bnbeta=0.1 * torch.randn(kn, device="cuda", generator=g),
    )
    return d


def fp64_error(layout, spec, d, nsamp=512):
    """max|y - y64| / max|y64| at sampled outputs (FP64 oracle, oracle/sampled.py)."""
    from oracle import Shape, sampled
    b, h, w, c, kn, kh, kw, pad, st = spec
    s = Shape(b, h, w, c, kn, kh, kw, pad, pad, st, st, 1, 1)
    rng = np.random.default_rng(123)
    idx = np.stack([rng.integers(0, n, nsamp) for n in (s.b, s.ho, s.wo, s.kn)], 1)
    host = {k: v.cpu().numpy() for k, v in d.items()}
    ref = sampled.fwd_sampled(layout, s, host["W"], host["X"], idx, bias=host["bias"], mean=host["mean"],
                              inv_std=host["inv_std"], gamma=host["gamma"], bnbeta=host["bnbeta"], relu=True)
    got = sampled.nhwc_index(layout, host["Y"], idx)
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-30))


def precision_sweep(torch, cg, stream, tune=True, reps=10, batch=64):
    """Forward layers at batch 64, NHWC: FP32 vs BF16 vs FP8-e4m3 operands
    (per-tensor amax scaling, FP32 accumulation), X = |N(0,1)| (post-ReLU),
    W ~ N(0, 0.02).  Per layer: TFLOP/s of the convolution kernels (operands
    already converted; the conversion is timed separately) and the error
    against the FP64 oracle at sampled outputs.  Plans: the loaded tuning
    database, or autotuned here (untimed) for problems it lacks (all of them
    with tune=True and no database)."""
    from oracle import Shape, sampled
    res = {}
    for li, (name, spec) in enumerate(LAYERS):
        _, h, w, c, kn, kh, kw, pad, st = spec
        spec = (batch,) + spec[1:]
        s = Shape(batch, h, w, c, kn, kh, kw, pad, pad, st, st, 1, 1)
        g = torch.Generator(device="cuda").manual_seed(0xC0FFEE + 4000 + li)
        X = torch.randn(s.x_shape(0), device="cuda", generator=g).abs()
        W = torch.randn(s.w_shape(0), device="cuda", generator=g) * 0.02
        Y = torch.empty(s.y_shape(0), device="cuda")
        desc = cg.make_desc(cg.FWD, 0, batch, h, w, c, kn, kh, kw, pad, pad, st, st)
        flops = layer_flops(spec)
        rng = np.random.default_rng(li)
        idx = np.stack([rng.integers(0, n, 384) for n in (s.b, s.ho, s.wo, s.kn)], 1)
        ref = sampled.fwd_sampled(0, s, W.cpu().numpy(), X.cpu().numpy(), idx)
        row = {}
        for dt, name_dt in ((cg.F32, "fp32"), (cg.BF16, "bf16"), (cg.FP8_E4M3, "fp8")):
            if dt == cg.F32:
                Xo, Wo, scales, conv_ms = X, W, None, 0.0
            else:
                tdt = torch.int16 if dt == cg.BF16 else torch.uint8
                Xo = torch.empty(X.shape, dtype=tdt, device="cuda")
                Wo = torch.empty(W.shape, dtype=tdt, device="cuda")
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                sx = cg.quantize(X, Xo, dt, stream=stream)
                sw = cg.quantize(W, Wo, dt, stream=stream)
                e1.record(stream)
                torch.cuda.synchronize()
                conv_ms = e0.elapsed_time(e1)
                scales = (sw, sx) if dt == cg.FP8_E4M3 else None
            if tune and not cg.tune_get(desc, dt):
                cg.autotune(desc, Wo, Xo, Y, dtype=dt, scales=scales, reps=3, stream=stream)
            for _ in range(3):
                cg.conv_ex(desc, Wo, Xo, Y, dtype=dt, scales=scales, stream=stream)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                cg.conv_ex(desc, Wo, Xo, Y, dtype=dt, scales=scales, stream=stream)
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            got = sampled.nhwc_index(0, Y.cpu().numpy(), idx)
            err = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
            rel2 = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
            row[name_dt] = {"ms": round(ms, 4), "tflops": round(flops / (ms * 1e-3) / 1e12, 2),
                            "roofline": roofline_of(flops, op_bytes(spec, "fwd", name_dt), name_dt, ms),
                            "fp64_max_rel_err": err, "fp64_l2_rel_err": rel2,
                            "quantize_ms": round(conv_ms, 4), "plan": list(cg.last_plan()[:2])}
        for k in ("bf16", "fp8"):
            row[k]["speedup_vs_fp32"] = round(row["fp32"]["ms"] / row[k]["ms"], 2)
        res[name + "_b64_NHWC"] = row
    return res


def backward_sweep(torch, cg, stream, reps=10, batch=32):
    """Backward of the layer sweep at batch 32, NHWC: weight gradient
    dW = im2row(X)^T dY and backward data dX += row2im(W^T dY), FP32 and BF16
    operands (FP32 accumulation), plans from the loaded tuning database or
    autotuned at setup; TFLOP/s of the convolution
    kernels and the error against the FP64 oracle at sampled outputs (BF16
    against the unquantized FP32 data)."""
    from oracle import Shape, sampled
    res = {}
    for li, (name, spec) in enumerate(LAYERS):
        _, h, w, c, kn, kh, kw, pad, st = spec
        s = Shape(batch, h, w, c, kn, kh, kw, pad, pad, st, st, 1, 1)
        flops = layer_flops((batch,) + spec[1:])
        g = torch.Generator(device="cuda").manual_seed(0xC0FFEE + 3000 + li)
        X = torch.randn(s.x_shape(0), device="cuda", generator=g)
        W = torch.randn(s.w_shape(0), device="cuda", generator=g) * (2.0 / s.K) ** 0.5
        dY = torch.randn(s.y_shape(0), device="cuda", generator=g)
        rng = np.random.default_rng(li)
        ws = np.stack([rng.integers(0, n, 96) for n in (s.c, s.kh, s.kw, s.kn)], 1)
        xs = np.stack([rng.integers(0, n, 96) for n in (s.b, s.h, s.w, s.c)], 1)
        Xh, Wh, dYh = X.cpu().numpy(), W.cpu().numpy(), dY.cpu().numpy()
        ref_w = sampled.wgrad_sampled(0, s, dYh, Xh, ws)
        ref_x = sampled.dgrad_sampled(0, s, Wh, dYh, xs)
        row = {}
        for dt, dname in ((cg.F32, "fp32"), (cg.BF16, "bf16")):
            if dt == cg.F32:
                Xo, Wo, dYo = X, W, dY
            else:
                Xo, Wo, dYo = (torch.empty(t.shape, dtype=torch.int16, device="cuda") for t in (X, W, dY))
                for src, dst in ((X, Xo), (W, Wo), (dY, dYo)):
                    cg.quantize(src, dst, cg.BF16, stream=stream)
            dW = torch.empty(s.w_shape(0), device="cuda")
            dX = torch.zeros(s.x_shape(0), device="cuda")
            dw_desc = cg.make_desc(cg.WGRAD, 0, batch, h, w, c, kn, kh, kw, pad, pad, st, st)
            dx_desc = cg.make_desc(cg.DGRAD, 0, batch, h, w, c, kn, kh, kw, pad, pad, st, st)
            out = {}
            for op, desc, args in (("wgrad", dw_desc, (dYo, Xo, dW)), ("dgrad", dx_desc, (Wo, dYo, dX))):
                if not cg.tune_get(desc, dt):  # not in a loaded database: autotune (untimed)
                    cg.autotune(desc, *args, dtype=dt, reps=3, stream=stream)
                for _ in range(2):
                    cg.conv_ex(desc, *args, beta=1.0 if op == "dgrad" else 0.0, dtype=dt, stream=stream)
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for _ in range(reps):
                    cg.conv_ex(desc, *args, beta=1.0 if op == "dgrad" else 0.0, dtype=dt, stream=stream)
                e1.record(stream)
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / reps
                out[op] = {"ms": round(ms, 4), "tflops": round(flops / (ms * 1e-3) / 1e12, 2),
                           "roofline": roofline_of(flops, op_bytes((batch,) + spec[1:], op, dname), dname, ms),
                           "plan": list(cg.last_plan())}  # cfg, splits, kernels per call
            # errors of one clean evaluation of each
            cg.conv_ex(dw_desc, dYo, Xo, dW, dtype=dt, stream=stream)
            dX.zero_()
            cg.conv_ex(dx_desc, Wo, dYo, dX, beta=1.0, dtype=dt, stream=stream)
            torch.cuda.synchronize()
            gw = sampled.w_index(0, dW.cpu().numpy(), ws)
            gx = sampled.nhwc_index(0, dX.cpu().numpy(), xs)
            out["wgrad"]["fp64_max_rel_err"] = float(np.max(np.abs(gw - ref_w)) / np.max(np.abs(ref_w)))
            out["dgrad"]["fp64_max_rel_err"] = float(np.max(np.abs(gx - ref_x)) / np.max(np.abs(ref_x)))
            row[dname] = out
        for op in ("wgrad", "dgrad"):
            row["bf16"][op]["speedup_vs_fp32"] = round(row["fp32"][op]["ms"] / row["bf16"][op]["ms"], 2)
        res[name + "_b32_NHWC"] = row
    return res


def batch_sweep(torch, cg, stream, n=256, reps=10, seed=5):
    """Many uneven BF16 forward layers (NHWC; batch {1,4,16}, H=W {7..56},
    Cin/Cout 32..1024, k {1,3,5}, s {1,2}): one conv_ex call per problem vs
    one batch (grouped launches captured in a hipGraph; bounded: at most 16
    grouped GEMM launches, 18 kernels per run; single: every problem in ONE
    persistent launch, with the weight repack in the run or, with static
    weights, done once at creation), plans from the
    learned cost model shipped in convgemm_amd/data (trained on autotuner
    timings).  Returns aggregate TFLOP/s of each way of launching, and the
    learned and analytic plans' regret against exhaustive autotuning of the
    same problems."""
    from convgemm_amd import tuning
    model = os.path.join(ROOT, "convgemm_amd", "data", "mi355x_costmodel.mlp")
    cg.costmodel_load(model)
    cg.set_planner(cg.PLAN_MODEL)
    try:
        probs = tuning.random_problems(n, seed)
        g = torch.Generator(device="cuda").manual_seed(0xC0FFEE + 5000)
        ops, flops = [], 0.0
        for p in probs:
            d = tuning.desc_of(p)
            ho = out_hw(p["h"], p["pad"], p["k"], p["s"])
            x = torch.randn(p["b"] * p["h"] * p["h"] * p["c"], device="cuda", generator=g)
            w = torch.randn(p["c"] * p["k"] * p["k"] * p["kn"], device="cuda", generator=g) * 0.05
            y = torch.empty(p["b"] * ho * ho * p["kn"], device="cuda")
            ops.append((d, w.to(torch.bfloat16).view(torch.int16), x.to(torch.bfloat16).view(torch.int16), y))
            flops += 2.0 * p["b"] * ho * ho * p["c"] * p["k"] * p["k"] * p["kn"]

        def timed(fn):
            for _ in range(2):
                fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                fn()
            e1.record(stream)
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) / reps

        def sequential():
            for d, w, x, y in ops:
                cg.conv_ex(d, w, x, y, dtype=cg.BF16, stream=stream)

        res = {"problems": n, "dtype": "bf16", "gflop": round(flops / 1e9, 3), "planner": "learned cost model"}
        ms = timed(sequential)
        res["sequential_calls"] = {"ms": round(ms, 4), "tflops": round(flops / (ms * 1e-3) / 1e12, 2)}
        for key, fl in (("batch_graph", cg.BATCH_GRAPH),
                        ("batch_graph_persistent", cg.BATCH_GRAPH | cg.BATCH_PERSISTENT),
                        ("batch_graph_bounded", cg.BATCH_GRAPH | cg.BATCH_BOUNDED),
                        ("batch_graph_bounded_static_weights",
                         cg.BATCH_GRAPH | cg.BATCH_BOUNDED | cg.BATCH_STATIC_WEIGHTS),
                        ("batch_single", cg.BATCH_GRAPH | cg.BATCH_SINGLE),
                        ("batch_single_static_weights", cg.BATCH_GRAPH | cg.BATCH_SINGLE | cg.BATCH_STATIC_WEIGHTS)):
            b = cg.Batch([cg.problem(d, w, x, y) for d, w, x, y in ops], cg.BF16, fl)
            ms = timed(lambda: b.run(stream))
            gr, kern, _ = b.info()
            res[key] = {"ms": round(ms, 4), "tflops": round(flops / (ms * 1e-3) / 1e12, 2), "groups": gr,
                        "kernels_per_run": kern}
            b.close()
        res["speedup_graph_vs_sequential"] = round(res["sequential_calls"]["ms"] / res["batch_graph"]["ms"], 2)
        # exhaustive autotuning of the same problems (untimed) judges the learned plans:
        # regret t(chosen) / t(best) from the recorded timings, and the batch again on
        # the measured-best plans
        t0 = time.time()
        for d, w, x, y in ops:
            cg.autotune(d, w, x, y, dtype=cg.BF16, reps=3, stream=stream)
        res["autotune_seconds"] = round(time.time() - t0, 1)
        res["regret_learned_vs_exhaustive"] = tuning.regret(probs, cg.BF16, cg.PLAN_MODEL)
        res["regret_analytic_vs_exhaustive"] = tuning.regret(probs, cg.BF16, cg.PLAN_ANALYTIC)
        cg.set_planner(cg.PLAN_TUNED)
        b = cg.Batch([cg.problem(d, w, x, y) for d, w, x, y in ops], cg.BF16, cg.BATCH_GRAPH)
        ms = timed(lambda: b.run(stream))
        res["batch_graph_exhaustive_plans"] = {"ms": round(ms, 4), "tflops": round(flops / (ms * 1e-3) / 1e12, 2)}
        b.close()
        return res
    finally:
        cg.set_planner(cg.PLAN_TUNED)


def explicit_vs_implicit(torch, cg, stream, reps=10, batch=32):
    """The reference drivers' ablation (tests/test_base.c:45-88) on the GPU, per
    layer at batch 32, FP32 and BF16 operands: the explicit path -- zero-fill
    the patch matrix, im2row (NHWC) / im2col (NCHW) into it, a plain GEMM on
    it (the same MFMA kernels, run as a 1x1 convolution over the patch
    matrix), the bias / batch-norm / ReLU epilogue as separate
    elementwise passes (torch, like the reference's separate loops), and for
    NCHW the transpose to [b][kn][ho][wo] -- against the implicit call with
    the fused epilogue.  Every stage is timed alone with HIP events over
    `reps` back-to-back launches; the explicit total is the sum of its
    stages.  Both results are compared (max|explicit - implicit| / max|.|)."""
    res = {}
    for li, (name, spec) in enumerate(LAYERS):
        b, h, w, c, kn, kh, kw, pad, st = spec
        ho, wo = out_hw(h, pad, kh, st), out_hw(w, pad, kw, st)
        P, K = batch * ho * wo, c * kh * kw
        flops = layer_flops((batch,) + spec[1:])
        for layout in (0, 1):
            d = make_inputs(torch, layout, (batch,) + spec[1:], 0xC0FFEE + 6000 + 10 * li + layout)
            desc = cg.make_desc(cg.FWD, layout, batch, h, w, c, kn, kh, kw, pad, pad, st, st)
            for dt, dname in ((cg.F32, "fp32"), (cg.BF16, "bf16")):
                if dt == cg.F32:
                    X, W, tdt = d["X"], d["W"], torch.float32
                else:
                    X = torch.empty(d["X"].shape, dtype=torch.int16, device="cuda")
                    W = torch.empty(d["W"].shape, dtype=torch.int16, device="cuda")
                    cg.quantize(d["X"], X, cg.BF16, stream=stream)
                    cg.quantize(d["W"], W, cg.BF16, stream=stream)
                    tdt = torch.int16
                pm = torch.empty(P * K, dtype=tdt, device="cuda")  # patch matrix: rows P x K / cols K x P
                Yi = torch.empty(d["Y"].shape, device="cuda")
                Ye = torch.empty(d["Y"].shape, device="cuda")
                Yg = torch.empty(kn * P, device="cuda") if layout == 1 else Ye
                vec = (lambda v: v.view(1, kn, 1, 1)) if layout == 1 else (lambda v: v)
                bias, mean, inv_std = vec(d["bias"]), vec(d["mean"]), vec(d["inv_std"])
                gamma, bnb = vec(d["gamma"]), vec(d["bnbeta"])

                def zero():
                    pm.zero_()

                def transform():
                    cg.im2row_ex(desc, X, pm, K if layout == 0 else P, dtype=dt, stream=stream)

                # the plain GEMM on the patch matrix, as a 1x1 convolution on the same MFMA kernels:
                # NHWC Y (P x kn) = rows (P x K) W (K x kn); NCHW Yg (kn x P) = W (kn x K) cols (K x P)
                gdesc = (cg.make_desc(cg.FWD, 0, P, 1, 1, K, kn, 1, 1) if layout == 0
                         else cg.make_desc(cg.FWD, 1, 1, 1, P, K, kn, 1, 1))

                def gemm():
                    cg.conv_ex(gdesc, W, pm, Ye if layout == 0 else Yg, dtype=dt, stream=stream)

                def transpose():
                    cg.transpose_nchw_ex(P, kn, Yg, P, 0.0, Ye, kn, ho, wo, stream=stream)

                def epilogue():
                    Ye.add_(bias)
                    Ye.sub_(mean).mul_(inv_std).mul_(gamma).add_(bnb)
                    Ye.clamp_(min=0.0)

                def implicit():
                    cg.conv_ex(desc, W, X, Yi, bias=d["bias"], mean=d["mean"], inv_std=d["inv_std"],
                               gamma=d["gamma"], bnbeta=d["bnbeta"], relu=True, dtype=dt, stream=stream)

                def timed(fn):
                    fn()
                    torch.cuda.synchronize()
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(stream)
                    for _ in range(reps):
                        fn()
                    e1.record(stream)
                    torch.cuda.synchronize()
                    return e0.elapsed_time(e1) / reps

                with torch.cuda.stream(stream):
                    stages = {"zero_fill": timed(zero), "im2row" if layout == 0 else "im2col": timed(transform),
                              "gemm": timed(gemm)}
                    if layout == 1:
                        stages["transpose_nchw"] = timed(transpose)
                    stages["epilogue_passes"] = timed(epilogue)
                    t_imp = timed(implicit)
                    # one clean explicit evaluation, compared with the implicit result
                    zero(); transform(); gemm()
                    if layout == 1:
                        transpose()
                    epilogue()
                    torch.cuda.synchronize()
                diff = float((Ye - Yi).abs().max() / Yi.abs().max().clamp_min(1e-30))
                t_exp = sum(stages.values())
                res[f"{name}_{'NHWC' if layout == 0 else 'NCHW'}_{dname}"] = {
                    "implicit": {"ms": round(t_imp, 4), "tflops": round(flops / (t_imp * 1e-3) / 1e12, 2)},
                    "explicit": {"ms": round(t_exp, 4), "tflops": round(flops / (t_exp * 1e-3) / 1e12, 2),
                                 "stages_ms": {k: round(v, 4) for k, v in stages.items()},
                                 "patch_matrix_bytes": P * K * (4 if dt == cg.F32 else 2)},
                    "speedup_implicit": round(t_exp / t_imp, 2),
                    "transform_share_of_explicit": round(
                        (stages["zero_fill"] + stages.get("im2row", 0) + stages.get("im2col", 0)
                         + stages.get("transpose_nchw", 0)) / t_exp, 3),
                    "max_rel_diff_explicit_vs_implicit": diff}
                del pm, Yi, Ye, Yg
    return res


REF_CATALOGUE_MEDIANS_GFS = {  # FP32, batch 64, host "volta" (plots/test.out, test_back.out, test_trans.out)
    "fwd_nhwc": 906.0, "fwd_nchw": 437.0, "dgrad_nhwc": 366.0, "dgrad_nchw": 159.0,
    "wgrad_nhwc": 296.0, "wgrad_nchw": 293.0}


def catalogue_shapes(batch=64):
    """The reference's CNN layer catalogue (tests/test.txt, committed as
    tests/golden/cnn_layer_shapes.txt; columns h w c b kh kw kn stride padding):
    the distinct rows at `batch`, as tests/test.pl feeds them to the drivers --
    220 shapes at batch 64, the set plots/test*.out hold."""
    seen, rows = set(), []
    for line in open(os.path.join(ROOT, "tests", "golden", "cnn_layer_shapes.txt")):
        if line.startswith("#") or not line.strip():
            continue
        h, w, c, b, kh, kw, kn, st, pad = map(int, line.split())
        key = (h, w, c, b, kh, kw, kn, st, pad)
        if b != batch or key in seen:
            continue
        seen.add(key)
        rows.append((b, h, w, c, kn, kh, kw, pad, st))
    return sorted(rows, key=lambda s: s[1] * s[2] * s[3] * s[0] * s[5] * s[6] * s[4])


def catalogue_sweep(torch, cg, stream, reps=3, batch=64):
    """FP32 throughput over the reference's whole shape catalogue at its batch
    (220 layers, batch 64), the set its published logs hold: forward with the
    fused bias + batch-norm + ReLU epilogue, weight gradient and backward data
    (dX += ...), NHWC and NCHW, each planned by the default planner (the
    learned cost model; no autotuning), timed with HIP events over `reps`
    back-to-back calls after one warm-up call.  GF/s = 2*m*n*k / t as in
    plots/convgemm.plt:56-57.  Reported: the median and quartiles over the
    catalogue per operation beside the reference's own medians on its CPU,
    and, as a size-independent check of every shape, the adjoint identity
    <conv(X,W), dY> = <X, dgrad(W,dY)> = <W, wgrad(X,dY)> (plain calls, FP64
    dot products on the GPU)."""
    shapes = catalogue_shapes(batch)
    gfs = {k: [] for k in REF_CATALOGUE_MEDIANS_GFS}
    worst_gap = 0.0
    worst = None
    t_start = time.perf_counter()
    for si, spec in enumerate(shapes):
        b, h, w, c, kn, kh, kw, pad, st = spec
        ho, wo = out_hw(h, pad, kh, st), out_hw(w, pad, kw, st)
        if ho <= 0 or wo <= 0:
            continue
        flops = layer_flops(spec)
        for layout, lname in ((0, "nhwc"), (1, "nchw")):
            d = make_inputs(torch, layout, spec, 0xC0FFEE + 9000 + 2 * si + layout)
            dY = torch.randn(d["Y"].shape, device="cuda")
            dX = torch.zeros(d["X"].shape, device="cuda")
            dW = torch.empty(d["W"].shape, device="cuda")
            fdesc = cg.make_desc(cg.FWD, layout, b, h, w, c, kn, kh, kw, pad, pad, st, st)
            wdesc = cg.make_desc(cg.WGRAD, layout, b, h, w, c, kn, kh, kw, pad, pad, st, st)
            ddesc = cg.make_desc(cg.DGRAD, layout, b, h, w, c, kn, kh, kw, pad, pad, st, st)
            calls = {
                "fwd": lambda: cg.conv_ex(fdesc, d["W"], d["X"], d["Y"], bias=d["bias"], mean=d["mean"],
                                          inv_std=d["inv_std"], gamma=d["gamma"], bnbeta=d["bnbeta"], relu=True,
                                          stream=stream),
                "wgrad": lambda: cg.conv_ex(wdesc, dY, d["X"], dW, stream=stream),
                "dgrad": lambda: cg.conv_ex(ddesc, d["W"], dY, dX, beta=1.0, stream=stream),
            }
            for op, fn in calls.items():
                fn()
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for _ in range(reps):
                    fn()
                e1.record(stream)
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / reps
                gfs[f"{op}_{lname}"].append(flops / (ms * 1e-3) / 1e9)
            # adjoint identity on plain calls (no epilogue, dX overwritten with alpha=1, beta=0)
            cg.conv_ex(fdesc, d["W"], d["X"], d["Y"], stream=stream)
            cg.conv_ex(wdesc, dY, d["X"], dW, stream=stream)
            cg.conv_ex(ddesc, d["W"], dY, dX, beta=0.0, stream=stream)
            torch.cuda.synchronize()
            s_y = float((d["Y"].double() * dY.double()).sum())
            s_x = float((d["X"].double() * dX.double()).sum())
            s_w = float((d["W"].double() * dW.double()).sum())
            scale = max(float(d["Y"].double().norm() * dY.double().norm()), 1e-30)
            gap = max(abs(s_y - s_x), abs(s_y - s_w)) / scale
            if gap > worst_gap:
                worst_gap, worst = gap, list(spec)
            del d, dY, dX, dW
    out = {"shapes": len(shapes), "batch": batch, "dtype": "f32", "planner": "learned cost model (default)",
           "reps": reps, "wall_s": round(time.perf_counter() - t_start, 1)}
    for k, v in gfs.items():
        a = np.array(v)
        med = float(np.median(a))
        out[k] = {"median_gflops": round(med, 1), "q1": round(float(np.percentile(a, 25)), 1),
                  "q3": round(float(np.percentile(a, 75)), 1), "max": round(float(a.max()), 1),
                  "reference_cpu_median_gflops": REF_CATALOGUE_MEDIANS_GFS[k],
                  "ratio_vs_reference_median": round(med / REF_CATALOGUE_MEDIANS_GFS[k], 1)}
    out["adjoint_identity_max_rel_gap"] = worst_gap
    out["adjoint_identity_worst_shape"] = worst
    return out


def _cpu_problems(batch=32):
    """The bench step's ten forward convolutions at `batch` on the host (FP32, seeded)."""
    from oracle import Shape
    rng = np.random.default_rng(7)
    probs = []
    for layout in (0, 1):
        for name, spec in LAYERS:
            _, h, w, c, kn, kh, kw, pad, st = spec
            s = Shape(batch, h, w, c, kn, kh, kw, pad, pad, st, st, 1, 1)
            X = rng.standard_normal(s.x_shape(layout), dtype=np.float32)
            W = (rng.standard_normal(s.w_shape(layout), dtype=np.float32) * np.float32(np.sqrt(2.0 / s.K)))
            v = lambda: rng.standard_normal(kn, dtype=np.float32)  # noqa: E731
            epi = dict(bias=0.1 * v(), mean=0.1 * v(), inv_std=(1.0 / np.sqrt(rng.uniform(0.5, 2.0, kn) + 1e-5)).astype(
                np.float32), gamma=rng.uniform(0.5, 1.5, kn).astype(np.float32), bnbeta=0.1 * v(), relu=True)
            probs.append((layout, s, W, X, epi, np.empty(s.y_shape(layout), np.float32)))
    return probs


def cpu_worker(reps):
    """One process of the CPU baseline: 1 warm-up step, then `reps` timed steps of the
    oracle's FP32 restatement of the reference's implicit GEMM; prints the step times."""
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    threads = max(1, min(threads, len(os.sched_getaffinity(0))))
    oracle.port_set_threads(threads)
    probs = _cpu_problems()
    times = []
    for r in range(reps + 1):
        t0 = time.perf_counter()
        for layout, s, W, X, epi, out in probs:
            oracle.port_conv(layout, "N", s, W, X, out=out, **epi)
        if r > 0:
            times.append(time.perf_counter() - t0)
    print(json.dumps({"step_seconds": times, "threads": threads}), flush=True)
    return 0


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(procs=5, reps=5):
    """The bench step (the same ten batch-32 forward convolutions with the fused
    epilogue) on the host cores through the oracle's FP32 restatement of the
    reference's B3A2C0 implicit GEMM (oracle/convgemm_port.c: OpenMP, portable
    micro-kernel; the reference itself needs BLIS, which this image lacks).
    Protocol of the reference's drivers (tests/CMakeLists.txt:20,
    tests/test_base.c:44): `procs` separate processes, each 1 warm-up and
    `reps` timed steps; value = step FLOPs / median step time."""
    import subprocess
    times, threads = [], None
    for _ in range(procs):
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-worker", str(reps)],
                           stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"cpu baseline worker failed: {r.stderr[-2000:]}")
        rec = json.loads(r.stdout.strip().splitlines()[-1])
        times += rec["step_seconds"]
        threads = rec["threads"]
    step_flops = sum(layer_flops(spec) for _, spec in LAYERS) * 2
    med = float(np.median(times))
    return {"value": step_flops / med / 1e12, "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "cpu_model": _cpu_model(), "nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "median_step_s": round(med, 4), "step_s_min_max": [round(min(times), 4), round(max(times), 4)],
            "sample": f"the full bench step (10 forward convolutions, batch 32, fused bias+BN+ReLU): "
                      f"{procs} processes x (1 warm-up + {reps} timed steps), median of {len(times)}; FP32 "
                      f"restatement of the reference's B3A2C0 implicit GEMM (portable micro-kernel, OpenMP "
                      f"{threads} threads)"}


def copy_probe(torch, mib, steps, warmup):
    """Plain device-to-device copy of `mib` MiB: a kernel whose HBM traffic is known
    (reads and writes mib MiB each), to scale the FETCH_SIZE/WRITE_SIZE counters."""
    n = mib * (1 << 20) // 4
    src = torch.randn(n, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    print(json.dumps({"copy_probe_bytes": n * 4, "steps": steps, "ms": round(ms, 4),
                      "GBps_read_plus_write": round(2 * n * 4 / (ms * 1e-3) / 1e9, 1)}), flush=True)
    return 0


def vendor_gemm(torch, reps):
    """hipBLASLt (torch.matmul) on the GEMM each forward layer is at batch 64
    (M = b*ho*wo, K = c*kh*kw, N = kn, A materialized, no gather) and on a
    square 8192^3 problem: what the vendor library reaches on these shapes,
    next to which the implicit-GEMM convolution's TF/s can be read."""
    shapes = [(name, 64 * out_hw(h, pad, kh, st) * out_hw(w, pad, kw, st), c * kh * kw, kn)
              for name, (b, h, w, c, kn, kh, kw, pad, st) in LAYERS] + [("square", 8192, 8192, 8192)]
    out = {}
    for dt, tag in ((torch.bfloat16, "bf16"), (torch.float32, "f32")):
        for name, M, K, N in shapes:
            a = torch.randn(M, K, device="cuda").to(dt)
            b = torch.randn(K, N, device="cuda").to(dt)
            c = torch.empty(M, N, device="cuda", dtype=dt)
            for _ in range(3):
                torch.matmul(a, b, out=c)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                torch.matmul(a, b, out=c)
            e1.record()
            torch.cuda.synchronize()
            us = e0.elapsed_time(e1) * 1e3 / reps
            out[f"{name}_{tag}"] = {"M": M, "K": K, "N": N, "us": round(us, 2),
                                    "tflops": round(2.0 * M * K * N / us / 1e6, 1)}
            del a, b, c
    print(json.dumps({"vendor_gemm": "hipBLASLt via torch.matmul", "layers_at_batch": 64, "reps": reps,
                      "results": out}), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tune-reps", type=int, default=3, help="timed repetitions per candidate when autotuning")
    ap.add_argument("--no-error", action="store_true", help="skip the FP64 error check")
    ap.add_argument("--no-tune", action="store_true", help="use the analytic planner instead of autotuning at setup")
    ap.add_argument("--no-precision", action="store_true", help="skip the BF16/FP8 precision sweep")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched uneven-layer sweep")
    ap.add_argument("--no-backward", action="store_true", help="skip the backward (wgrad/dgrad) sweep")
    ap.add_argument("--tune-db", default="", help="tuning database: loaded if present (layers found there are "
                    "not re-tuned), written back after tuning")
    ap.add_argument("--fresh-tune", action="store_true",
                    help="do not load the shipped tuning database: autotune every layer at setup")
    ap.add_argument("--vendor-gemm", action="store_true",
                    help="only time hipBLASLt (torch.matmul) on the layers' GEMM shapes, for calibration")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB per step (calibrates HBM byte counters)")
    ap.add_argument("--precision-only", action="store_true",
                    help="only run the BF16/FP8 precision sweep and print it (for profiling its kernels)")
    ap.add_argument("--backward-only", action="store_true",
                    help="only run the backward sweep and print it (for profiling its kernels)")
    ap.add_argument("--no-ablation", action="store_true", help="skip the explicit-vs-implicit comparison")
    ap.add_argument("--no-catalogue", action="store_true",
                    help="skip the FP32 sweep over the reference's 220-layer shape catalogue at batch 64")
    ap.add_argument("--catalogue-only", action="store_true",
                    help="only run the shape-catalogue sweep and print it")
    ap.add_argument("--cpu-worker", type=int, default=0, metavar="REPS",
                    help="internal: run one CPU-baseline process with REPS timed steps and exit")
    ap.add_argument("--only", default="", help="comma-separated layer names (e.g. L2_3x3_64to64_56_NHWC) "
                    "to run alone, for profiling; the JSON value then covers only those")
    a = ap.parse_args(argv)
    if a.cpu_worker:  # a host-only child process: never touches the GPU
        return cpu_worker(a.cpu_worker)
    if a.gpus != 1:
        print(f"bench.py: this build targets one GPU; --gpus {a.gpus} is not supported "
              "(sharding the convolution by batch across GPUs is a design, not part of this build)", file=sys.stderr)
        return 2

    import torch
    import convgemm_amd as cg

    assert torch.cuda.is_available(), "bench.py needs an MI355X"
    cg.lib()
    stream = torch.cuda.current_stream()
    if a.copy_probe:
        return copy_probe(torch, a.copy_probe, a.steps, a.warmup)
    if a.vendor_gemm:
        return vendor_gemm(torch, a.steps)
    plans = "autotuned at setup"
    if a.tune_db and os.path.exists(a.tune_db):
        cg.tune_load(a.tune_db)
        plans = f"tuning database {os.path.basename(a.tune_db)}, then autotuned"
    elif not a.tune_db and not a.fresh_tune and not a.no_tune and os.path.exists(SHIPPED_TUNE_DB):
        cg.tune_load(SHIPPED_TUNE_DB)
        for db in (SHIPPED_BWD_DB, SHIPPED_PREC_DB):
            if os.path.exists(db):
                cg.tune_load(db)
        plans = "shipped tuning database (MI355X autotuner timings)"
    if a.catalogue_only:
        print(json.dumps({"catalogue_sweep": catalogue_sweep(torch, cg, stream)}), flush=True)
        return 0
    if a.precision_only:
        res = {"precision_sweep": precision_sweep(torch, cg, stream, tune=not a.no_tune), "plans": plans}
        if a.tune_db:
            cg.tune_save(a.tune_db)
        print(json.dumps(res), flush=True)
        return 0
    if a.backward_only:
        res = {"backward_sweep": backward_sweep(torch, cg, stream), "plans": plans}
        if a.tune_db:
            cg.tune_save(a.tune_db)
        print(json.dumps(res), flush=True)
        return 0
    convs = []
    for layout in (0, 1):
        for li, (name, spec) in enumerate(LAYERS):
            b, h, w, c, kn, kh, kw, pad, st = spec
            d = make_inputs(torch, layout, spec, 0xC0FFEE + 2000 + 10 * li + layout)
            desc = cg.make_desc(cg.FWD, layout, b, h, w, c, kn, kh, kw, pad, pad, st, st)
            convs.append(dict(name=f"{name}_{'NHWC' if layout == 0 else 'NCHW'}", layout=layout, spec=spec,
                              desc=desc, d=d, flops=layer_flops(spec), bytes=layer_bytes(spec)))

    if a.only:
        keep = set(a.only.split(","))
        convs = [cv for cv in convs if cv["name"] in keep]
        assert convs, f"--only matched nothing: {a.only}"

    def run(cv):
        d = cv["d"]
        cg.conv_ex(cv["desc"], d["W"], d["X"], d["Y"], bias=d["bias"], mean=d["mean"], inv_std=d["inv_std"],
                   gamma=d["gamma"], bnbeta=d["bnbeta"], relu=True, stream=stream)

    # plans: the autotuner times every tile configuration / split-K candidate
    # once per layer during setup (untimed), like a benchmark-mode library
    tuned = 0
    for cv in convs:
        if not a.no_tune and not cg.tune_get(cv["desc"]):
            d = cv["d"]
            cg.autotune(cv["desc"], d["W"], d["X"], d["Y"], reps=a.tune_reps, stream=stream)
            tuned += 1
    if a.tune_db and tuned:
        cg.tune_save(a.tune_db)
    if a.no_tune:
        cg.set_planner(cg.PLAN_ANALYTIC)
    for cv in convs:
        run(cv)
        cv["plan"] = cg.last_plan()
    for _ in range(a.warmup):
        for cv in convs:
            run(cv)
    torch.cuda.synchronize()

    # the timed region: exactly K steps of the ten calls, nothing else on the
    # stream (a timing event between calls costs several microseconds of idle
    # GPU on this platform, so the per-layer breakdown is a separate pass)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(a.steps):
        for cv in convs:
            run(cv)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # per-layer breakdown (untimed pass): HIP events around each call
    ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in convs]
          for _ in range(a.steps)]
    for k in range(a.steps):
        for i, cv in enumerate(convs):
            ev[k][i][0].record(stream)
            run(cv)
            ev[k][i][1].record(stream)
    torch.cuda.synchronize()

    step_flops = sum(cv["flops"] for cv in convs)
    elapsed = t1 - t0
    value = step_flops * a.steps / elapsed / 1e12
    per = []
    for i, cv in enumerate(convs):
        ms = float(np.mean([ev[k][i][0].elapsed_time(ev[k][i][1]) for k in range(a.steps)]))
        per.append(ms)
        cv["ms"] = ms
    # Dominant kernel: the implicit-GEMM kernel with the largest share of the
    # step, among layers whose GEMM is one kernel.  A call may also repack the
    # weights (FP32/BF16/FP8 forward reads [kn][kh][kw][c]); the kernel is
    # timed alone by launching it on weights packed once (conv_packed), with
    # HIP events on the launch stream, right after the timed steps.
    for cv in convs:
        d = cv["d"]
        psz = cg.packed_weights_size(cv["desc"])
        cv["packed"] = torch.empty(max(psz // 4, 1), device="cuda") if psz else d["W"]
        if psz:
            cg.pack_weights(cv["desc"], d["W"], cv["packed"], stream=stream)
        cg.conv_packed(cv["desc"], cv["packed"], d["X"], d["Y"], bias=d["bias"], mean=d["mean"],
                       inv_std=d["inv_std"], gamma=d["gamma"], bnbeta=d["bnbeta"], relu=True, stream=stream)
        cv["kplan"] = cg.last_plan()
    single = [cv for cv in convs if cv["kplan"][2] == 1] or convs
    dom = max(single, key=lambda cv: cv["ms"])
    dd = dom["d"]
    for _ in range(3):
        cg.conv_packed(dom["desc"], dom["packed"], dd["X"], dd["Y"], bias=dd["bias"], mean=dd["mean"],
                       inv_std=dd["inv_std"], gamma=dd["gamma"], bnbeta=dd["bnbeta"], relu=True, stream=stream)
    # K launches back to back between two events: their mean duration (an event
    # between launches would add its own latency to every sample)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(a.steps):
        cg.conv_packed(dom["desc"], dom["packed"], dd["X"], dd["Y"], bias=dd["bias"], mean=dd["mean"],
                       inv_std=dd["inv_std"], gamma=dd["gamma"], bnbeta=dd["bnbeta"], relu=True, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    kernel_ms = e0.elapsed_time(e1) / a.steps
    achieved = dom["flops"] / (kernel_ms * 1e-3) / 1e12
    # HBM bytes per launch of the dominant kernel: FETCH_SIZE/WRITE_SIZE counter
    # passes scaled by a plain copy (profiles/traffic_fwd_fp32.md), valid while
    # the layer runs the plan that was measured
    traffic = None
    traffic_source = "no committed counter figure for this kernel"
    tpath = os.path.join(ROOT, "profiles", "traffic_fwd_fp32.json")
    if os.path.exists(tpath):
        rec = json.load(open(tpath))["layers"].get(dom["name"])
        if rec and (rec["plan"]["cfg"], rec["plan"]["splits"]) == tuple(dom["plan"][:2]):
            traffic = rec["traffic_bytes"]
            traffic_source = ("profiles/traffic_fwd_fp32.json: rocprofv3 FETCH_SIZE and WRITE_SIZE of this "
                              "kernel's plan, scaled by a plain copy's readings (committed, not measured live)")
        elif rec:
            traffic_source = (f"the committed counter figure is for plan {rec['plan']}, this run used "
                              f"{list(dom['plan'][:2])}")

    errors = {}
    if not a.no_error:
        for cv in convs:
            errors[cv["name"]] = fp64_error(cv["layout"], cv["spec"], cv["d"])

    layers = {cv["name"]: {"ms": round(cv["ms"], 4), "tflops": round(cv["flops"] / (cv["ms"] * 1e-3) / 1e12, 2),
                           "frac_fp32_peak": round(cv["flops"] / (cv["ms"] * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 3),
                           "plan": {"cfg": cv["plan"][0], "splits": cv["plan"][1], "kernels": cv["plan"][2]},
                           "fp64_max_rel_err": errors.get(cv["name"])}
              for cv in convs}
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(elapsed / a.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "resnet50_layer_sweep_fwd_b32_nhwc_nchw_bias_bn_relu",
                   "layers": [n for n, _ in LAYERS], "layouts": ["NHWC", "NCHW"], "global_batch": 32,
                   "gflop_per_step": round(step_flops / 1e9, 3)},
        "plans": plans if not a.no_tune else "analytic planner",
        "roofline": {"bound": "mfma", "kernel": dom["name"], "kernel_us": round(kernel_ms * 1e3, 2),
                     "achieved": round(achieved, 3),
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "traffic": traffic, "traffic_source": traffic_source},
        "fp64_max_rel_err": max(errors.values()) if errors else None,
        "layers": layers,
    }
    if not a.no_precision:
        out["precision_sweep"] = precision_sweep(torch, cg, stream, tune=not a.no_tune)
    if not a.no_backward:
        out["backward_sweep"] = backward_sweep(torch, cg, stream)
    if not a.no_batch:
        out["batch_sweep"] = batch_sweep(torch, cg, stream)
    if not a.no_ablation:
        out["explicit_vs_implicit"] = explicit_vs_implicit(torch, cg, stream)
    if not a.no_catalogue:
        out["catalogue_sweep"] = catalogue_sweep(torch, cg, stream)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
