"""Full-size GPU checks of the reduced-precision and batched workloads.

* Backward of the ResNet-50 layer sweep at batch 32 in BF16 (weight gradient
  dW = im2row(X)^T dY and backward data dX += row2im(W^T dY), NHWC and NCHW),
  from BF16-quantized X ~ N(0,1), He-scaled W and dY ~ N(0,1).
* Forward of the same layers at batch 64 in BF16 and FP8-e4m3 (per-tensor
  amax scaling), X = |N(0,1)| (post-ReLU), W ~ N(0, 0.02).
* 256 uneven BF16 forward problems (batch {1,4,16}, H=W {7..56}, channels
  32..1024, kernels {1,3,5}, strides {1,2}) run as one graph batch on plans
  from the learned cost model.

Two bars per result, at sampled outputs of the FP64 oracle
(oracle/sampled.py, pinned against the direct oracle in test_oracle.py),
max-normwise (max|got - ref| / max|ref|):
* against the oracle on the DEQUANTIZED operands -- the kernel's own error,
  FP32 accumulation only: BF16 <= 2e-5, FP8 <= 1e-4;
* against the oracle on the UNQUANTIZED FP32 operands -- what the caller
  sees: BF16 <= 1e-2 (the target's bar); FP8 <= 8e-2 (e4m3 keeps 3 mantissa
  bits, a relative rounding error up to 2^-4 per operand).
The batch is also checked bit-identical, problem by problem, to one
convgemm_conv_ex call per problem with the same plans (a stream-K plan runs
in a batch as its configuration's whole tiles, so its reference call is made
with split-K 1).
"""
import os

import numpy as np
import pytest
import torch

from oracle import NCHW, NHWC, Shape, sampled

pytestmark = pytest.mark.gpu

import convgemm_amd as cg  # noqa: E402

LAYERS = {
    "L1_1x1_256to64_56": (56, 256, 64, 1, 1),
    "L2_3x3_64to64_56": (56, 64, 64, 3, 1),
    "L3_3x3s2_128to128_56": (56, 128, 128, 3, 2),
    "L4_1x1_1024to256_14": (14, 1024, 256, 1, 1),
    "L5_3x3_512to512_7": (7, 512, 512, 3, 1),
}
TOL_DEQ = {cg.BF16: 2e-5, cg.FP8_E4M3: 1e-4}
TOL_RAW = {cg.BF16: 1e-2, cg.FP8_E4M3: 8e-2}


def _shape(name, b):
    h, c, kn, k, s = LAYERS[name]
    return Shape(b, h, h, c, kn, k, k, k // 2, k // 2, s, s, 1, 1)


def _desc(op, layout, s):
    return cg.make_desc(op, layout, s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil,
                        s.hdil)


def _quant(t, dtype):
    """(stored operand, dequantized FP32 copy on the host, scale) of FP32 tensor t."""
    if dtype == cg.BF16:
        q = torch.empty(t.shape, dtype=torch.int16, device="cuda")
        cg.quantize(t, q, cg.BF16)
        torch.cuda.synchronize()
        return q, q.view(torch.bfloat16).float().cpu().numpy(), 1.0
    q = torch.empty(t.shape, dtype=torch.uint8, device="cuda")
    sc = cg.quantize(t, q, cg.FP8_E4M3)
    torch.cuda.synchronize()
    return q, (q.view(torch.float8_e4m3fn).float() * sc).cpu().numpy(), sc


def _err(got, ref):
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-30))


@pytest.mark.parametrize("layout", [NHWC, NCHW])
@pytest.mark.parametrize("name", list(LAYERS))
def test_bf16_backward_fullsize(name, layout):
    s = _shape(name, 32)
    print(f"bf16 backward {name} b32 layout={layout}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(0xC0FFEE + 3000 + sum(map(ord, name)) + 7 * layout)
    X = torch.randn(s.x_shape(layout), device="cuda", generator=g)
    W = torch.randn(s.w_shape(layout), device="cuda", generator=g) * (2.0 / s.K) ** 0.5
    dY = torch.randn(s.y_shape(layout), device="cuda", generator=g)
    (Xq, Xd, _), (Wq, Wd, _), (Yq, Yd, _) = (_quant(t, cg.BF16) for t in (X, W, dY))
    Xh, Wh, dYh = X.cpu().numpy(), W.cpu().numpy(), dY.cpu().numpy()
    dW = torch.full(s.w_shape(layout), float("nan"), device="cuda")
    cg.conv_ex(_desc(cg.WGRAD, layout, s), Yq, Xq, dW, dtype=cg.BF16)
    dX = torch.zeros(s.x_shape(layout), device="cuda")
    cg.conv_ex(_desc(cg.DGRAD, layout, s), Wq, Yq, dX, beta=1.0, dtype=cg.BF16)
    torch.cuda.synchronize()
    rng = np.random.default_rng(31 + layout)
    ws = np.stack([rng.integers(0, n, 128) for n in (s.c, s.kh, s.kw, s.kn)], 1)
    got = sampled.w_index(layout, dW.cpu().numpy(), ws)
    e_deq = _err(got, sampled.wgrad_sampled(layout, s, Yd, Xd, ws))
    e_raw = _err(got, sampled.wgrad_sampled(layout, s, dYh, Xh, ws))
    print(f"  wgrad: vs dequantized {e_deq:.2e}, vs FP32 operands {e_raw:.2e}", flush=True)
    assert e_deq <= TOL_DEQ[cg.BF16] and e_raw <= TOL_RAW[cg.BF16], (e_deq, e_raw)
    xs = np.stack([rng.integers(0, n, 128) for n in (s.b, s.h, s.w, s.c)], 1)
    got = sampled.nhwc_index(layout, dX.cpu().numpy(), xs)
    e_deq = _err(got, sampled.dgrad_sampled(layout, s, Wd, Yd, xs))
    e_raw = _err(got, sampled.dgrad_sampled(layout, s, Wh, dYh, xs))
    print(f"  dgrad: vs dequantized {e_deq:.2e}, vs FP32 operands {e_raw:.2e}", flush=True)
    assert e_deq <= TOL_DEQ[cg.BF16] and e_raw <= TOL_RAW[cg.BF16], (e_deq, e_raw)


@pytest.mark.parametrize("dtype", [cg.BF16, cg.FP8_E4M3])
@pytest.mark.parametrize("layout", [NHWC, NCHW])
@pytest.mark.parametrize("name", list(LAYERS))
def test_lowp_forward_fullsize_b64(name, layout, dtype):
    s = _shape(name, 64)
    print(f"forward {name} b64 layout={layout} dtype={dtype}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(0xC0FFEE + 4000 + sum(map(ord, name)) + 7 * layout)
    X = torch.randn(s.x_shape(layout), device="cuda", generator=g).abs()
    W = torch.randn(s.w_shape(layout), device="cuda", generator=g) * 0.02
    (Xq, Xd, sx), (Wq, Wd, sw) = _quant(X, dtype), _quant(W, dtype)
    Y = torch.full(s.y_shape(layout), float("nan"), device="cuda")
    scales = (sw, sx) if dtype == cg.FP8_E4M3 else None
    cg.conv_ex(_desc(cg.FWD, layout, s), Wq, Xq, Y, dtype=dtype, scales=scales)
    torch.cuda.synchronize()
    rng = np.random.default_rng(41 + layout)
    ys = np.stack([rng.integers(0, n, 128) for n in (s.b, s.ho, s.wo, s.kn)], 1)
    got = sampled.nhwc_index(layout, Y.cpu().numpy(), ys)
    e_deq = _err(got, sampled.fwd_sampled(layout, s, Wd, Xd, ys))
    e_raw = _err(got, sampled.fwd_sampled(layout, s, W.cpu().numpy(), X.cpu().numpy(), ys))
    print(f"  plan {cg.last_plan()}: vs dequantized {e_deq:.2e}, vs FP32 operands {e_raw:.2e}", flush=True)
    assert e_deq <= TOL_DEQ[dtype] and e_raw <= TOL_RAW[dtype], (e_deq, e_raw)


def test_uneven_batch_256_learned_plans():
    from convgemm_amd import tuning
    model = os.path.join(os.path.dirname(cg.__file__), "data", "mi355x_costmodel.mlp")
    cg.costmodel_load(model)
    prev = cg.get_planner()
    cg.set_planner(cg.PLAN_MODEL)
    try:
        probs = tuning.random_problems(256, seed=5)
        gen = torch.Generator(device="cuda").manual_seed(0xC0FFEE + 5000)
        cases = []
        for p in probs:
            d = tuning.desc_of(p)
            ho = (p["h"] + 2 * p["pad"] - p["k"]) // p["s"] + 1
            x = torch.randn(p["b"], p["h"], p["h"], p["c"], device="cuda", generator=gen).abs()
            w = torch.randn(p["c"], p["k"], p["k"], p["kn"], device="cuda", generator=gen) * 0.05
            cases.append(dict(p=p, d=d, x=x.to(torch.bfloat16), w=w.to(torch.bfloat16),
                              y=torch.full((p["b"], ho, ho, p["kn"]), float("nan"), device="cuda")))
        print("per-call references", flush=True)
        refs = []
        for c in cases:
            r = torch.empty_like(c["y"])
            cg.conv_ex(c["d"], c["w"].view(torch.int16), c["x"].view(torch.int16), r, dtype=cg.BF16)
            cfg, splits = cg.last_plan()[:2]
            if splits == 0:  # a batch runs a stream-K plan as its configuration's whole tiles
                cg.force_config(cfg, 1)
                try:
                    cg.conv_ex(c["d"], c["w"].view(torch.int16), c["x"].view(torch.int16), r, dtype=cg.BF16)
                finally:
                    cg.force_config(-1, 0)
            refs.append(r)
        torch.cuda.synchronize()
        print("graph batch", flush=True)
        b = cg.Batch([cg.problem(c["d"], c["w"].view(torch.int16), c["x"].view(torch.int16), c["y"])
                      for c in cases], cg.BF16, cg.BATCH_GRAPH)
        for rep in range(2):  # replay twice, eager work between
            for c in cases:
                c["y"].fill_(float("nan"))
            torch.cuda.synchronize()
            b.run()
            torch.cuda.synchronize()
            for i, (c, r) in enumerate(zip(cases, refs)):
                assert torch.equal(c["y"], r), (rep, i, c["p"])
        b.close()
        # the largest problem and 15 others against the FP64 oracle on the BF16 operands
        fl = [p["b"] * p["c"] * p["kn"] * p["k"] ** 2 * ((p["h"] + 2 * p["pad"] - p["k"]) // p["s"] + 1) ** 2
              for p in probs]
        order = [int(np.argmax(fl))] + [i for i in range(0, 256, 17) if i != int(np.argmax(fl))][:15]

        def oracle_check(tag):
            rng = np.random.default_rng(51)
            for i in order:
                c, p = cases[i], cases[i]["p"]
                s = Shape(p["b"], p["h"], p["h"], p["c"], p["kn"], p["k"], p["k"], p["pad"], p["pad"], p["s"],
                          p["s"], 1, 1)
                ys = np.stack([rng.integers(0, n, 64) for n in (s.b, s.ho, s.wo, s.kn)], 1)
                ref = sampled.fwd_sampled(NHWC, s, c["w"].float().cpu().numpy(), c["x"].float().cpu().numpy(), ys)
                got = sampled.nhwc_index(NHWC, c["y"].cpu().numpy(), ys)
                e = _err(got, ref)
                print(f"  {tag} problem {i} {p}: vs FP64 on the BF16 operands {e:.2e}", flush=True)
                assert e <= TOL_DEQ[cg.BF16], (tag, i, p, e)

        oracle_check("graph batch")
        # the same problems in ONE persistent launch (static weights, graph): one kernel per run,
        # each output bit-identical to the per-call run of the batch's plan
        print("single-launch batch", flush=True)
        b = cg.Batch([cg.problem(c["d"], c["w"].view(torch.int16), c["x"].view(torch.int16), c["y"])
                      for c in cases], cg.BF16, cg.BATCH_GRAPH | cg.BATCH_STATIC_WEIGHTS | cg.BATCH_SINGLE)
        assert b.info() == (1, 1, True), b.info()
        refs = []
        for i, c in enumerate(cases):
            cfg, splits = b.plan(i)
            r = torch.empty_like(c["y"])
            cg.force_config(cfg, splits)
            try:
                cg.conv_ex(c["d"], c["w"].view(torch.int16), c["x"].view(torch.int16), r, dtype=cg.BF16)
            finally:
                cg.force_config(-1, 0)
            refs.append(r)
        torch.cuda.synchronize()
        for rep in range(2):
            for c in cases:
                c["y"].fill_(float("nan"))
            torch.cuda.synchronize()
            b.run()
            torch.cuda.synchronize()
            for i, (c, r) in enumerate(zip(cases, refs)):
                assert torch.equal(c["y"], r), (rep, i, c["p"], b.plan(i))
        b.close()
        oracle_check("single launch")
    finally:
        cg.set_planner(prev)
