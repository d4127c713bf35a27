"""Pipelined weight-gradient kernel (NHWC; FP32 and BF16 operands, FP32
accumulation): pixel rows loaded straight into LDS, MFMA operands read
transposed (BF16: ds_read_b64_tr_b16).

Every configuration of the family is forced, with split-K 1, 2 and 3, on
shapes with channel / kn counts that are not multiples of the tile, pixel
counts that are not multiples of the K step, strides, asymmetric padding,
dilation and 1x1 / 5x5 kernels, and checked against the FP64 oracle (BF16 on
the dequantized operands, where every product is exact in FP32 and only the
accumulation rounds): max-rel <= 1e-5 (FP32) and <= 2e-5 (BF16).  alpha and
beta (dW = alpha * im2row(X)^T dY + beta * dW) are checked the same way, and
grouped launches (batches, persistent workgroups, graph replays) must be
bit-identical to single calls.
"""
import numpy as np
import pytest
import torch

import oracle
from oracle import NHWC, Shape

pytestmark = pytest.mark.gpu

import convgemm_amd as cg  # noqa: E402
from tests.gpu_util import DEV, rel_err  # noqa: E402
from tests.test_gpu_batch import Case, D, check  # noqa: E402
from tests.test_gpu_lowp import q_bf16  # noqa: E402

SHAPES = [
    Shape(2, 12, 11, 32, 48, 3, 3, 1, 1, 1, 1, 1, 1),    # C < BM, kn < BN, P = 264 (ragged K step)
    Shape(2, 9, 9, 24, 40, 3, 3, 1, 0, 2, 2, 1, 1),      # stride 2, asymmetric padding
    Shape(1, 14, 14, 136, 72, 1, 1, 0, 0, 1, 1, 1, 1),   # 1x1, C = 128 + 8 (two M tiles per tap)
    Shape(2, 10, 10, 16, 24, 3, 3, 2, 2, 1, 1, 2, 2),    # dilation 2
    Shape(3, 8, 8, 40, 16, 5, 5, 2, 2, 1, 1, 1, 1),      # 5x5
    Shape(4, 28, 28, 64, 64, 3, 3, 1, 1, 1, 1, 1, 1),    # many K steps (P = 3136)
    Shape(1, 16, 16, 264, 200, 3, 3, 1, 1, 1, 1, 1, 1),  # ragged C and kn beyond 256
]

DTYPES = {"f32": cg.F32, "bf16": cg.BF16}


def lpw_configs(elem):
    """Raster-order-0 configurations of the pipelined weight-gradient family."""
    return [i for i, t in enumerate(cg.tile_configs(elem))
            if t[5] > 0 and t[6] == 0 and cg.tile_supports(elem, i, cg.WGRAD, NHWC)
            and not cg.tile_supports(elem, i, cg.FWD, NHWC)]


def operands(dtype, X, dY):
    if dtype == "f32":
        return torch.from_numpy(X).to(DEV), X, torch.from_numpy(dY).to(DEV), dY
    (Xd, Xq), (dYd, dYq) = q_bf16(X), q_bf16(dY)
    return Xd, Xq, dYd, dYq


def test_family_exists():
    for elem in DTYPES.values():
        cfgs = lpw_configs(elem)
        assert len(cfgs) >= 6, cfgs
        for c in cfgs:  # weight gradient NHWC only, 16-byte groups only
            assert not cg.tile_supports(elem, c, cg.WGRAD, cg.NCHW)
            assert not cg.tile_supports(elem, c, cg.WGRAD, NHWC, vec=False)
            assert not cg.tile_supports(elem, c, cg.DGRAD, NHWC)
    assert not lpw_configs(cg.FP8_E4M3)


@pytest.mark.parametrize("dtype", list(DTYPES))
@pytest.mark.parametrize("si", range(len(SHAPES)))
def test_every_configuration(dtype, si):
    elem = DTYPES[dtype]
    s = SHAPES[si]
    rng = np.random.default_rng(700 + si)
    X = rng.standard_normal(s.x_shape(NHWC)).astype(np.float32)
    dY = rng.standard_normal(s.y_shape(NHWC)).astype(np.float32)
    Xd, Xq, dYd, dYq = operands(dtype, X, dY)
    ref = oracle.wgrad(NHWC, s, dYq, Xq)
    tol = 1e-5 if dtype == "f32" else 2e-5
    geo = (s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil, s.hdil)
    d = cg.make_desc(cg.WGRAD, NHWC, *geo)
    for cfg in lpw_configs(elem):
        for splits in (1, 2, 3):
            print(f"lpw {dtype} shape {si} cfg {cfg} splits {splits}", flush=True)
            cg.force_config(cfg, splits)
            try:
                dW = torch.full(s.w_shape(NHWC), float("nan"), device=DEV)
                cg.conv_ex(d, dYd, Xd, dW, dtype=elem)
                torch.cuda.synchronize()
                assert cg.last_plan()[0] == cfg
            finally:
                cg.force_config(-1, 0)
            assert rel_err(dW.cpu().numpy(), ref)[0] <= tol, (cfg, splits)


@pytest.mark.parametrize("dtype", list(DTYPES))
def test_alpha_beta(dtype):
    """dW = alpha * im2row(X)^T dY + beta * dW_old, direct and split-K write-back."""
    elem = DTYPES[dtype]
    s = SHAPES[0]
    rng = np.random.default_rng(77)
    X = rng.standard_normal(s.x_shape(NHWC)).astype(np.float32)
    dY = rng.standard_normal(s.y_shape(NHWC)).astype(np.float32)
    old = rng.standard_normal(s.w_shape(NHWC)).astype(np.float32)
    Xd, Xq, dYd, dYq = operands(dtype, X, dY)
    ref = 0.5 * oracle.wgrad(NHWC, s, dYq, Xq) + 2.0 * old.astype(np.float64)
    geo = (s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil, s.hdil)
    d = cg.make_desc(cg.WGRAD, NHWC, *geo)
    for cfg in lpw_configs(elem)[:2]:
        for splits in (1, 3):
            cg.force_config(cfg, splits)
            try:
                dW = torch.from_numpy(old).to(DEV)
                cg.conv_ex(d, dYd, Xd, dW, alpha=0.5, beta=2.0, dtype=elem)
                torch.cuda.synchronize()
            finally:
                cg.force_config(-1, 0)
            assert rel_err(dW.cpu().numpy(), ref)[0] <= 1e-5, (cfg, splits)


@pytest.mark.parametrize("dtype", [cg.F32, cg.BF16])
@pytest.mark.parametrize("flags", [cg.BATCH_GRAPH, cg.BATCH_GRAPH | cg.BATCH_PERSISTENT])
def test_grouped_bit_identical(dtype, flags):
    """Batches of weight-gradient problems on each configuration of the family
    (grouped, persistent) give the single calls' bits, over two graph replays."""
    W = cg.WGRAD
    for cfg in lpw_configs(dtype):
        for splits in (1, 2):
            print(f"lpw group dtype {dtype} cfg {cfg} splits {splits}", flush=True)
            cg.force_config(cfg, splits)
            try:
                cases = [Case(D(W, NHWC, 2, 10, 10, 24, 40, 3), dtype, 31 + cfg),
                         Case(D(W, NHWC, 1, 9, 11, 32, 16, 3, s=2), dtype, 32 + cfg, beta=1.0),
                         Case(D(W, NHWC, 3, 7, 7, 72, 64, 1), dtype, 33 + cfg, alpha=0.5),
                         Case(D(W, NHWC, 1, 12, 12, 16, 24, 5), dtype, 34 + cfg)]
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
