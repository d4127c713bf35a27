"""Backward data of unit-stride NHWC convolutions in its forward form.

dX is the forward convolution of dY with the kernel turned by 180 degrees
(taps reversed, c and kn swapped, padding (K - 1) * dilation - pad), so these
problems run on the forward kernels (pipelined and halo tiles).  Checks, FP32
and BF16 operands:
* against the FP64 oracle (dgrad), with beta accumulation, dilation,
  asymmetric padding, 1x1 / 3x3 / 5x5 kernels and channel tails;
* against the residue-class kernel (the hook that switches the forward form
  off), within the accumulation tolerance;
* padding beyond the kernel's reach keeps the residue-class kernel;
* a batch of such problems (its repacks reverse the taps in the grouped
  element-wise launch) is bit-identical to one call per problem;
* full-size L2 and L5 at batch 32 against sampled FP64 outputs.
"""
import numpy as np
import pytest
import torch

import oracle
from oracle import NHWC, Shape, sampled

pytestmark = pytest.mark.gpu

import convgemm_amd as cg  # noqa: E402
from tests.gpu_util import DEV, rel_err  # noqa: E402
from tests.test_gpu_lowp import q_bf16  # noqa: E402

SHAPES = [
    Shape(2, 12, 11, 32, 48, 3, 3, 1, 1, 1, 1, 1, 1),
    Shape(2, 10, 9, 24, 40, 3, 3, 2, 1, 1, 1, 2, 2),   # dilation 2, asymmetric padding
    Shape(1, 9, 13, 16, 64, 5, 5, 2, 1, 1, 1, 1, 1),   # 5x5, padding below the kernel's reach
    Shape(3, 7, 7, 64, 32, 1, 1, 0, 0, 1, 1, 1, 1),    # 1x1: no repack at all
    Shape(2, 8, 8, 20, 12, 3, 3, 1, 1, 1, 1, 1, 1),    # kn not a multiple of the 16-byte chunk
    Shape(2, 6, 30, 40, 24, 3, 3, 0, 1, 1, 1, 1, 1),   # no vertical padding, wide rows
]


def _run_dgrad(s, W, dY, dX0, dtype, beta):
    d = cg.make_desc(cg.DGRAD, NHWC, s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil,
                     s.hdil)
    out = dX0.clone()
    cg.conv_ex(d, W, dY, out, beta=beta, dtype=dtype)
    torch.cuda.synchronize()
    return out, cg.last_plan()


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
@pytest.mark.parametrize("si", range(len(SHAPES)))
def test_dgrad_forward_form(dtype, si):
    s = SHAPES[si]
    rng = np.random.default_rng(300 + si)
    W = (rng.standard_normal(s.w_shape(NHWC)) * 0.1).astype(np.float32)
    dY = rng.standard_normal(s.y_shape(NHWC)).astype(np.float32)
    dX0 = rng.standard_normal(s.x_shape(NHWC)).astype(np.float32)
    if dtype == "f32":
        Wd, Wq, dYd, dYq, elem, tol = torch.from_numpy(W).to(DEV), W, torch.from_numpy(dY).to(DEV), dY, cg.F32, 1e-5
    else:
        (Wd, Wq), (dYd, dYq), elem, tol = q_bf16(W), q_bf16(dY), cg.BF16, 2e-5
    X0 = torch.from_numpy(dX0).to(DEV)
    for beta in (0.0, 1.0):
        ref = oracle.dgrad(NHWC, s, Wq, dYq, beta=beta, dX0=dX0)
        print(f"dgrad forward form {dtype} shape {si} beta {beta}", flush=True)
        got, plan = _run_dgrad(s, Wd, dYd, X0, elem, beta)
        assert rel_err(got.cpu().numpy(), ref)[0] <= tol, (beta, plan)
        cg.set_dgrad_forward(False)
        try:
            old, plan_old = _run_dgrad(s, Wd, dYd, X0, elem, beta)
        finally:
            cg.set_dgrad_forward(True)
        assert rel_err(old.cpu().numpy(), ref)[0] <= tol
        scale = max(float(np.max(np.abs(ref))), 1.0)
        assert float((got - old).abs().max()) <= 4 * tol * scale
        # the forward form runs a forward configuration
        assert cg.tile_supports(elem, plan[0], cg.FWD, cg.NHWC)


def test_padding_beyond_reach_keeps_class_kernel():
    """pad > (K - 1) * dilation: the forward form would need negative padding."""
    s = Shape(2, 9, 9, 16, 16, 3, 3, 3, 3, 1, 1, 1, 1)
    rng = np.random.default_rng(5)
    W = (rng.standard_normal(s.w_shape(NHWC)) * 0.1).astype(np.float32)
    dY = rng.standard_normal(s.y_shape(NHWC)).astype(np.float32)
    ref = oracle.dgrad(NHWC, s, W, dY)
    got, plan = _run_dgrad(s, torch.from_numpy(W).to(DEV), torch.from_numpy(dY).to(DEV),
                           torch.zeros(s.x_shape(NHWC), device=DEV), cg.F32, 0.0)
    assert rel_err(got.cpu().numpy(), ref)[0] <= 1e-5
    cg.set_dgrad_forward(False)
    try:
        _, plan_old = _run_dgrad(s, torch.from_numpy(W).to(DEV), torch.from_numpy(dY).to(DEV),
                                 torch.zeros(s.x_shape(NHWC), device=DEV), cg.F32, 0.0)
    finally:
        cg.set_dgrad_forward(True)
    assert plan == plan_old


@pytest.mark.parametrize("dtype", [cg.F32, cg.BF16])
def test_dgrad_forward_batch(dtype):
    from tests.test_gpu_batch import Case, D, check
    rng = np.random.default_rng(61)
    cases = []
    for i in range(8):
        k = int(rng.choice([1, 3, 5]))
        cases.append(Case(D(cg.DGRAD, cg.NHWC, int(rng.choice([1, 2])), int(rng.choice([7, 12])),
                            int(rng.choice([7, 12])), int(rng.choice([16, 32, 48])), int(rng.choice([16, 32, 64])),
                            k), dtype, 700 + i, beta=1.0 if i % 3 == 2 else 0.0))
    refs = [c.reference() for c in cases]
    for flags in (cg.BATCH_GRAPH, cg.BATCH_GRAPH | cg.BATCH_STATIC_WEIGHTS):
        b = cg.Batch([c.problem() for c in cases], dtype, flags)
        for rep in range(2):
            for c in cases:
                c.reset()
            b.run()
            torch.cuda.synchronize()
            check(cases, refs, f"flags={flags} replay={rep}")
        b.close()


LAYERS = {
    "L2_3x3_64to64_56": Shape(32, 56, 56, 64, 64, 3, 3, 1, 1, 1, 1, 1, 1),
    "L5_3x3_512to512_7": Shape(32, 7, 7, 512, 512, 3, 3, 1, 1, 1, 1, 1, 1),
}


@pytest.mark.parametrize("name", list(LAYERS))
def test_dgrad_forward_fullsize(name):
    s = LAYERS[name]
    g = torch.Generator(device="cuda").manual_seed(sum(map(ord, name)) + 9)
    W = torch.randn(s.w_shape(NHWC), device="cuda", generator=g) * (2.0 / s.K) ** 0.5
    dY = torch.randn(s.y_shape(NHWC), device="cuda", generator=g)
    got, plan = _run_dgrad(s, W, dY, torch.zeros(s.x_shape(NHWC), device=DEV), cg.F32, 0.0)
    rng = np.random.default_rng(13)
    xs = np.stack([rng.integers(0, n, 256) for n in (s.b, s.h, s.w, s.c)], 1)
    ref = sampled.dgrad_sampled(NHWC, s, W.cpu().numpy(), dY.cpu().numpy(), xs)
    v = sampled.nhwc_index(NHWC, got.cpu().numpy(), xs)
    assert np.max(np.abs(v - ref)) <= 1e-5 * np.max(np.abs(ref)), plan
