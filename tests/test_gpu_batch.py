"""Batched grouped launches and hipGraph replay (include/convgemm_batch.h).

A batch runs the same kernels with the same plans as one convgemm_conv_ex
call per problem (a stream-K plan runs as its configuration's whole tiles),
so every output must be BIT-identical to the per-problem call of that plan:
the grouped kernel only changes which workgroup computes which tile, never
the arithmetic or its order.  Problems mix all six
operations, epilogues, alpha/beta, split-K, stride classes, BF16 and FP8.
Graph batches are replayed several times with eager work (a fill, a
conv_ex call) and synchronizations between replays, and every replay is
checked.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

import convgemm_amd as cg  # noqa: E402
from tests.gpu_util import DEV  # noqa: E402

FLAG_SETS = [0, cg.BATCH_GRAPH, cg.BATCH_PERSISTENT, cg.BATCH_GRAPH | cg.BATCH_PERSISTENT,
             cg.BATCH_GRAPH | cg.BATCH_STATIC_WEIGHTS]


def out_hw(h, pad, dil, k, s):
    return (h + 2 * pad - dil * (k - 1) - 1) // s + 1


def sizes(d):
    ho, wo = out_hw(d.h, d.vpad, d.vdil, d.kh, d.vstride), out_hw(d.w, d.hpad, d.hdil, d.kw, d.hstride)
    nx, nw, ny = d.b * d.h * d.w * d.c, d.c * d.kh * d.kw * d.kn, d.b * ho * wo * d.kn
    if d.op == cg.FWD:
        return nw, nx, ny  # in, x, out
    if d.op == cg.WGRAD:
        return ny, nx, nw
    return nw, ny, nx


def operand(n, dtype, gen):
    r = torch.randn(n, generator=gen, device="cpu").to(DEV)
    if dtype == cg.BF16:
        return r.to(torch.bfloat16).view(torch.int16)
    if dtype == cg.FP8_E4M3:
        return r.to(torch.float8_e4m3fn).view(torch.uint8)
    return r


class Case:
    """One problem with its operands, a pristine output and the per-call reference."""

    def __init__(self, desc, dtype, seed, alpha=1.0, beta=0.0, epi=None):
        g = torch.Generator().manual_seed(seed)
        ni, nx, no = sizes(desc)
        self.desc, self.dtype, self.alpha, self.beta = desc, dtype, alpha, beta
        self.in_, self.x = operand(ni, dtype, g), operand(nx, dtype, g)
        self.out0 = torch.randn(no, generator=g).to(DEV)
        self.out = self.out0.clone()
        self.scales = (0.5, 0.25) if dtype == cg.FP8_E4M3 else None
        self.epi = {}
        if epi and desc.op == cg.FWD:
            kn = desc.kn
            for k in epi:
                if k == "relu":
                    self.epi["relu"] = True
                elif k == "inv_std":
                    self.epi[k] = (torch.rand(kn, generator=g) + 0.5).to(DEV)
                else:
                    self.epi[k] = (0.1 * torch.randn(kn, generator=g)).to(DEV)

    def problem(self):
        return cg.problem(self.desc, self.in_, self.x, self.out, alpha=self.alpha, beta=self.beta,
                          scales=self.scales, **self.epi)

    def reference(self):
        """The per-call result of the plan a batch runs: the call's own plan, except that a
        stream-K plan runs in a batch as its configuration's whole tiles (split-K 1)."""
        ref = self.out0.clone()
        cg.conv_ex(self.desc, self.in_, self.x, ref, alpha=self.alpha, beta=self.beta, dtype=self.dtype,
                   scales=self.scales, **self.epi)
        cfg, splits = cg.last_plan()[:2]
        if splits == 0:  # only the planner picks stream-K here (no test forces splits 0 on a batch)
            ref.copy_(self.out0)
            cg.force_config(cfg, 1)
            try:
                cg.conv_ex(self.desc, self.in_, self.x, ref, alpha=self.alpha, beta=self.beta, dtype=self.dtype,
                           scales=self.scales, **self.epi)
            finally:
                cg.force_config(-1, 0)
        torch.cuda.synchronize()
        return ref

    def reset(self):
        self.out.copy_(self.out0)


def D(op, layout, b, h, w, c, kn, k, pad=None, s=1, dil=1):
    pad = k // 2 if pad is None else pad
    return cg.make_desc(op, layout, b, h, w, c, kn, k, k, pad, pad, s, s, dil, dil)


BN_RELU = ("bias", "mean", "inv_std", "gamma", "bnbeta", "relu")


def mixed_f32():
    F, W, G = cg.FWD, cg.WGRAD, cg.DGRAD
    return [
        Case(D(F, cg.NHWC, 2, 14, 14, 32, 48, 3), cg.F32, 1, epi=BN_RELU),
        Case(D(F, cg.NCHW, 1, 9, 11, 20, 36, 3, s=2), cg.F32, 2, epi=("bias", "relu")),
        Case(D(F, cg.NHWC, 4, 7, 7, 64, 64, 1), cg.F32, 3, alpha=0.5, beta=2.0),
        Case(D(F, cg.NHWC, 1, 28, 28, 16, 24, 5, s=2), cg.F32, 4),
        Case(D(W, cg.NHWC, 2, 10, 10, 24, 40, 3), cg.F32, 5),
        Case(D(W, cg.NCHW, 2, 12, 9, 16, 20, 3, s=2), cg.F32, 6, beta=1.0),
        Case(D(G, cg.NHWC, 2, 13, 13, 24, 32, 3, s=2), cg.F32, 7, beta=1.0),
        Case(D(G, cg.NCHW, 1, 10, 12, 18, 22, 3, dil=2), cg.F32, 8, beta=1.0),
        Case(D(F, cg.NCHW, 3, 16, 16, 40, 24, 1), cg.F32, 9, epi=("mean", "inv_std", "gamma", "bnbeta")),
        Case(D(F, cg.NHWC, 2, 8, 8, 16, 16, 3), cg.F32, 10, alpha=0.0, beta=3.0),  # beta scaling only
        Case(D(F, cg.NHWC, 0, 8, 8, 16, 16, 3), cg.F32, 11),  # empty batch: dropped
    ]


def check(cases, refs, tag):
    for i, (c, r) in enumerate(zip(cases, refs)):
        got, exp = c.out.cpu().numpy(), r.cpu().numpy()
        assert np.array_equal(got, exp), (tag, i, float(np.max(np.abs(got - exp))))


@pytest.mark.parametrize("flags", FLAG_SETS)
def test_mixed_f32_bit_identical(flags):
    cases = mixed_f32()
    refs = [c.reference() for c in cases]
    b = cg.Batch([c.problem() for c in cases], cg.F32, flags)
    groups, kernels, graph = b.info()
    assert groups >= 1 and kernels >= groups and graph == bool(flags & cg.BATCH_GRAPH)
    for rep in range(3 if flags & cg.BATCH_GRAPH else 1):
        for c in cases:
            c.reset()  # eager fills between replays
        torch.cuda.synchronize()
        b.run()
        torch.cuda.synchronize()
        check(cases, refs, f"flags={flags} replay={rep}")
        # unrelated eager library call between replays
        cases[0].reference()
    b.close()


@pytest.mark.parametrize("cfg,splits", [(3, 3), (0, 4), (6, 2)])
def test_split_k_groups(cfg, splits):
    cg.force_config(cfg, splits)
    try:
        cases = [c for c in mixed_f32() if c.desc.b > 0 and c.alpha != 0.0]
        refs = [c.reference() for c in cases]
        b = cg.Batch([c.problem() for c in cases], cg.F32, cg.BATCH_GRAPH)
    finally:
        cg.force_config(-1, 0)
    for rep in range(2):
        for c in cases:
            c.reset()
        b.run()
        torch.cuda.synchronize()
        check(cases, refs, f"cfg={cfg} splits={splits} replay={rep}")
    b.close()


def uneven(n, dtype, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        k = int(rng.choice([1, 3, 5]))
        s = int(rng.choice([1, 2]))
        h = int(rng.choice([7, 14, 28]))
        c = int(rng.choice(range(8, 129, 8)))
        kn = int(rng.choice(range(8, 129, 8)))
        lay = int(rng.integers(2))
        out.append(Case(D(cg.FWD, lay, int(rng.choice([1, 2, 4])), h, h, c, kn, k, s=s), dtype, 1000 + i,
                        epi=BN_RELU if i % 3 == 0 else None))
    return out


@pytest.mark.parametrize("dtype", [cg.F32, cg.BF16, cg.FP8_E4M3])
@pytest.mark.parametrize("flags", [cg.BATCH_GRAPH, cg.BATCH_GRAPH | cg.BATCH_PERSISTENT])
def test_uneven_forward_batch(dtype, flags):
    cases = uneven(40, dtype, 77 + dtype)
    refs = [c.reference() for c in cases]
    b = cg.Batch([c.problem() for c in cases], dtype, flags)
    for rep in range(2):
        for c in cases:
            c.reset()
        torch.cuda.synchronize()
        b.run(torch.cuda.current_stream())
        torch.cuda.synchronize()
        check(cases, refs, f"dtype={dtype} flags={flags} replay={rep}")
    b.close()


def pipelined_cases(dtype, layout, n, seed):
    """Forward problems the pipelined kernels serve: 16-byte channel chunks,
    padding, stride 2, ragged M/N/K tails, with and without the epilogue."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        k = int(rng.choice([1, 3]))
        s = int(rng.choice([1, 2]))
        h = int(rng.choice([7, 12, 20]))
        c = int(rng.choice(range(16, 161, 16)))
        kn = int(rng.choice(range(16, 161, 16)))
        out.append(Case(D(cg.FWD, layout, int(rng.choice([1, 2, 3])), h, h, c, kn, k, s=s), dtype, seed + i,
                        alpha=0.75 if i % 4 == 1 else 1.0, beta=1.0 if i % 4 == 3 else 0.0,
                        epi=BN_RELU if i % 2 == 0 else None))
    return out


@pytest.mark.parametrize("dtype,layout", [(cg.F32, cg.NHWC), (cg.F32, cg.NCHW), (cg.BF16, cg.NHWC),
                                          (cg.FP8_E4M3, cg.NHWC)])
@pytest.mark.parametrize("flags", [cg.BATCH_GRAPH, cg.BATCH_GRAPH | cg.BATCH_PERSISTENT])
def test_pipelined_groups(dtype, layout, flags):
    """Grouped (and persistent) launches of the pipelined kernels: every
    pipelined configuration forced in turn, with and without split-K."""
    cfgs = [i for i, t in enumerate(cg.tile_configs(dtype)) if t[5] > 0 and t[6] == 0 and t[7] == 0 and cg.tile_supports(dtype, i, cg.FWD, layout)]
    assert cfgs
    for cfg in cfgs:
        for splits in (1, 2):
            print(f"pipelined group dtype {dtype} layout {layout} cfg {cfg} splits {splits}", flush=True)
            cg.force_config(cfg, splits)
            try:
                cases = pipelined_cases(dtype, layout, 6, 400 + 10 * cfg + splits)
                refs = [c.reference() for c in cases]
                assert cg.last_plan()[0] == cfg
                b = cg.Batch([c.problem() for c in cases], dtype, flags)
            finally:
                cg.force_config(-1, 0)
            for rep in range(2):
                for c in cases:
                    c.reset()
                b.run()
                torch.cuda.synchronize()
                check(cases, refs, f"cfg={cfg} splits={splits} replay={rep}")
            b.close()


def test_bf16_backward_batch():
    F, W, G = cg.FWD, cg.WGRAD, cg.DGRAD
    cases = [Case(D(W, cg.NHWC, 2, 10, 10, 32, 48, 3), cg.BF16, 21),
             Case(D(G, cg.NHWC, 2, 11, 11, 40, 24, 3, s=2), cg.BF16, 22, beta=1.0),
             Case(D(W, cg.NCHW, 1, 9, 9, 24, 16, 3), cg.BF16, 23),
             Case(D(G, cg.NCHW, 2, 8, 10, 16, 32, 3), cg.BF16, 24, beta=1.0),
             Case(D(F, cg.NCHW, 2, 8, 8, 16, 32, 3), cg.BF16, 25, epi=BN_RELU)]
    refs = [c.reference() for c in cases]
    b = cg.Batch([c.problem() for c in cases], cg.BF16, cg.BATCH_GRAPH | cg.BATCH_STATIC_WEIGHTS)
    for c in cases:
        c.reset()
    b.run()
    torch.cuda.synchronize()
    check(cases, refs, "bf16 backward")
    b.close()


def test_batch_errors():
    c = Case(D(cg.FWD, cg.NHWC, 1, 8, 8, 8, 8, 3), cg.F32, 31)
    pr = c.problem()
    pr.out = None
    with pytest.raises(cg.ConvGemmError):
        cg.Batch([pr], cg.F32, 0)
    pr = c.problem()
    pr.desc.vstride = 0
    with pytest.raises(cg.ConvGemmError):
        cg.Batch([pr], cg.F32, 0)
    with pytest.raises(cg.ConvGemmError):
        cg.Batch([c.problem()], cg.F32, 64)  # unknown flag
    with pytest.raises(cg.ConvGemmError):  # FP8 is forward-only
        cg.Batch([Case(D(cg.WGRAD, cg.NHWC, 1, 8, 8, 16, 16, 3), cg.FP8_E4M3, 32).problem()], cg.FP8_E4M3, 0)
    empty = cg.Batch([], cg.F32, cg.BATCH_GRAPH)
    empty.run()
    torch.cuda.synchronize()
    assert empty.info()[0] == 0
    empty.close()


@pytest.mark.parametrize("dtype", [cg.F32, cg.BF16])
def test_bounded_batch(dtype, tmp_path):
    """CONVGEMM_BATCH_BOUNDED on a batch whose plans (set through a tuning
    database) name 24 kernel instances: at most 16 grouped launches, one
    element-wise launch and one reduction per run.  Problems that kept their
    plan are bit-identical to their per-call run; every output matches its
    per-call run within the accumulation tolerance."""
    ops = [(cg.FWD, cg.NHWC), (cg.FWD, cg.NCHW), (cg.WGRAD, cg.NHWC), (cg.WGRAD, cg.NCHW), (cg.DGRAD, cg.NHWC),
           (cg.DGRAD, cg.NCHW)]
    confs = cg.tile_configs(dtype)
    ns = cg.num_tile_shapes(dtype)
    rng = np.random.default_rng(77)
    cases, lines = [], []
    for i in range(24):
        op, layout = ops[i % 6]
        b, c, kn, st = int(rng.choice([1, 2])), int(rng.choice([16, 32, 48])), int(rng.choice([16, 24, 40])), \
            int(rng.choice([1, 2]))
        desc = D(op, layout, b, 10, 10, c, kn, 3, s=st)
        opk = 2 * [cg.FWD, cg.WGRAD, cg.DGRAD].index(op) + (layout == cg.NCHW)
        usable = [k for k in range(ns) if confs[k][7] == 0 and cg.tile_supports(dtype, k, op, layout)]
        cfg, splits = usable[(i // 6) % len(usable)], 1 + (i % 3 == 0)
        for vec in (0, 1):  # a plan for either load width the problem gets
            lines.append(f"{int(dtype)} {opk} {vec} {b} 10 10 {c} {kn} 3 3 1 1 {st} {st} 1 1 1 {cfg} {splits} 1.0")
        cases.append(Case(desc, dtype, 900 + i, beta=1.0 if i % 5 == 4 else 0.0,
                          epi=BN_RELU if op == cg.FWD else None))
    db = tmp_path / "plans.tunedb"
    db.write_text("\n".join(lines) + "\n")
    prev = cg.get_planner()
    cg.tune_clear()
    cg.tune_load(str(db))
    cg.set_planner(cg.PLAN_TUNED)
    try:
        refs, plans = [], set()
        for c in cases:
            refs.append(c.reference())
            plans.add((cg.last_plan()[0], c.desc.op, c.desc.layout))
        assert len(plans) > 16, plans
        b = cg.Batch([c.problem() for c in cases], dtype, cg.BATCH_GRAPH | cg.BATCH_BOUNDED)
    finally:
        cg.set_planner(prev)
        cg.tune_clear()
    groups, kernels, graph = b.info()
    assert groups <= 16 and kernels <= 18 and graph, (groups, kernels)
    for rep in range(2):
        for c in cases:
            c.reset()
        b.run()
        torch.cuda.synchronize()
        same = 0
        for i, (c, r) in enumerate(zip(cases, refs)):
            got, exp = c.out.cpu().numpy(), r.cpu().numpy()
            scale = max(float(np.max(np.abs(exp))), 1e-30)
            assert float(np.max(np.abs(got - exp))) <= 1e-5 * scale, (rep, i)
            same += bool(np.array_equal(got, exp))
        assert same >= len(cases) // 2, same
    b.close()


# ---- CONVGEMM_BATCH_SINGLE: forward NHWC problems in one persistent launch ----

CHUNK = {cg.F32: 4, cg.BF16: 8, cg.FP8_E4M3: 16}  # channels per 16-byte chunk


def single_cases(dtype, n, seed):
    """Forward NHWC problems with 16-byte channel chunks: ragged M/N/K tails,
    stride 2, 1x1 to 5x5, epilogues, alpha/beta."""
    rng = np.random.default_rng(seed)
    q = CHUNK[dtype]
    out = []
    for i in range(n):
        k = int(rng.choice([1, 3, 5]))
        s = int(rng.choice([1, 2]))
        h = int(rng.choice([7, 12, 20, 28]))
        c = q * int(rng.integers(1, 160 // q + 1))
        kn = int(rng.choice(range(16, 193, 8)))
        out.append(Case(D(cg.FWD, cg.NHWC, int(rng.choice([1, 2, 4])), h, h, c, kn, k, s=s), dtype, seed + i,
                        alpha=0.75 if i % 5 == 1 else 1.0, beta=1.0 if i % 5 == 3 else 0.0,
                        epi=BN_RELU if i % 2 == 0 else None))
    return out


def forced_reference(c, cfg, splits):
    """The per-call result of plan (cfg, splits)."""
    ref = c.out0.clone()
    cg.force_config(cfg, splits)
    try:
        cg.conv_ex(c.desc, c.in_, c.x, ref, alpha=c.alpha, beta=c.beta, dtype=c.dtype, scales=c.scales, **c.epi)
        assert cg.last_plan()[:2] == (cfg, splits)
    finally:
        cg.force_config(-1, 0)
    return ref


def run_and_check(b, cases, refs, tag, reps):
    for rep in range(reps):
        for c in cases:
            c.reset()  # eager fills between replays
        torch.cuda.synchronize()
        b.run()
        torch.cuda.synchronize()
        check(cases, refs, f"{tag} replay={rep}")


@pytest.mark.parametrize("dtype", [cg.F32, cg.BF16, cg.FP8_E4M3])
@pytest.mark.parametrize("flags", [cg.BATCH_SINGLE, cg.BATCH_SINGLE | cg.BATCH_GRAPH | cg.BATCH_STATIC_WEIGHTS])
def test_single_launch_batch(dtype, flags):
    """Every eligible problem on a tile shape of the single-launch kernel, split-K 1;
    with static weights and a graph the run is ONE kernel; outputs bit-identical to
    per-call runs of the batch's plans."""
    L = cg.lib()
    cases = single_cases(dtype, 40, 5000 + 100 * int(dtype))
    b = cg.Batch([c.problem() for c in cases], dtype, flags)
    groups, kernels, graph = b.info()
    plans = [b.plan(i) for i in range(len(cases))]
    for (cfg, splits), c in zip(plans, cases):
        assert L.cg_single_variant(int(dtype), cfg) >= 0 and splits == 1, (cfg, splits, c.desc)
    assert groups == 1
    if flags & cg.BATCH_STATIC_WEIGHTS:  # otherwise an element-wise launch repacks the weights per run
        assert kernels == 1, (groups, kernels, graph)
    refs = [forced_reference(c, cfg, splits) for c, (cfg, splits) in zip(cases, plans)]
    torch.cuda.synchronize()
    run_and_check(b, cases, refs, f"single dtype={dtype} flags={flags}", 3 if flags & cg.BATCH_GRAPH else 1)
    b.close()


@pytest.mark.parametrize("dtype", [cg.F32, cg.BF16, cg.FP8_E4M3])
def test_single_launch_every_shape(dtype):
    """Each tile shape of the single-launch kernel forced in turn (raster orders 0
    and 2), beside problems the kernel does not serve (NCHW, 8-byte channel rows,
    alpha = 0), which run in the usual grouped launches."""
    L = cg.lib()
    ns = cg.num_tile_shapes(dtype)
    shapes = [k for k in range(ns) if L.cg_single_variant(int(dtype), k) >= 0]
    assert len(shapes) >= 4
    q = CHUNK[dtype]
    for base in shapes:
        for raster in (0, 2):
            cfg = base + raster * ns
            print(f"single shape dtype {dtype} cfg {cfg}", flush=True)
            cases = single_cases(dtype, 6, 7000 + cfg)
            others = [Case(D(cg.FWD, cg.NCHW, 2, 9, 9, 16, 24, 3), dtype, 31, epi=BN_RELU),
                      Case(D(cg.FWD, cg.NHWC, 2, 8, 8, 16, 16, 3), dtype, 32, alpha=0.0, beta=2.0)]
            if dtype != cg.FP8_E4M3:
                others.append(Case(D(cg.FWD, cg.NHWC, 1, 10, 10, q + q // 2, 40, 3), dtype, 33))
            allc = cases + others
            cg.force_config(cfg, 1)
            try:
                b = cg.Batch([c.problem() for c in allc], dtype, cg.BATCH_SINGLE | cg.BATCH_GRAPH)
            finally:
                cg.force_config(-1, 0)
            plans = [b.plan(i) for i in range(len(allc))]
            assert all(p == (cfg, 1) for p in plans[:len(cases)]), plans
            assert plans[len(cases) + 1][0] == -1  # alpha = 0: a scaling only
            assert b.info()[0] >= 2  # the single launch and at least one grouped launch
            refs = [forced_reference(c, *p) if p[0] >= 0 else c.reference() for c, p in zip(allc, plans)]
            torch.cuda.synchronize()
            run_and_check(b, allc, refs, f"single cfg={cfg}", 2)
            b.close()
