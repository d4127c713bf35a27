"""BF16 and FP8-e4m3 operand paths (FP32 accumulation on MFMA).

* Quantizers: bit-exact against torch's CPU conversions (round to nearest
  even) -- FP32->BF16, and FP32->e4m3fn of x/scale with scale = amax/448.
* Convolutions: the HIP result must match the FP64 oracle evaluated on the
  *dequantized* operands (every product of two BF16 or FP8 values is exact in
  FP32, so the only error left is accumulation): max-rel <= 2e-5 for BF16 and
  <= 1e-4 for FP8 (the FP8 MFMA adds each 64-product block inside the matrix
  core before the FP32 accumulate, measured at ~2e-5).
* The error against FP64 on the unquantized data is the precision cost; it
  is checked against loose bounds here and reported by bench.py per layer
  (BF16 target <= 1e-2).
"""
import numpy as np
import pytest
import torch

import oracle
from oracle import NCHW, NHWC, Shape

pytestmark = pytest.mark.gpu

FP8_TOL = 1e-4

import convgemm_amd as cg  # noqa: E402
from tests.gpu_util import DEV, rel_err  # noqa: E402


def q_bf16(a_np):
    """FP32 numpy -> (device bf16 bits tensor, float64 dequantized numpy)."""
    src = torch.from_numpy(np.ascontiguousarray(a_np, np.float32)).to(DEV)
    dst = torch.empty(src.shape, dtype=torch.int16, device=DEV)
    assert cg.quantize(src, dst, cg.BF16) == 1.0
    deq = dst.cpu().view(torch.bfloat16).double().numpy()
    return dst, deq


def q_fp8(a_np):
    src = torch.from_numpy(np.ascontiguousarray(a_np, np.float32)).to(DEV)
    dst = torch.empty(src.shape, dtype=torch.uint8, device=DEV)
    scale = cg.quantize(src, dst, cg.FP8_E4M3)
    deq = dst.cpu().view(torch.float8_e4m3fn).double().numpy() * scale
    return dst, deq, scale


def test_bf16_quantizer_bit_exact():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.standard_normal(100000) * 10.0 ** rng.integers(-30, 30, 100000),
                        [0.0, -0.0, 1.0, -1.0, 3.4e38, -3.4e38, 1e-40, np.inf, -np.inf]]).astype(np.float32)
    src = torch.from_numpy(x).to(DEV)
    dst = torch.empty(x.shape, dtype=torch.int16, device=DEV)
    cg.quantize(src, dst, cg.BF16)
    ref = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16)
    np.testing.assert_array_equal(dst.cpu().numpy(), ref.numpy())


def test_fp8_quantizer_bit_exact_and_scale():
    rng = np.random.default_rng(2)
    x = (rng.standard_normal(200001) * 3.0).astype(np.float32)
    x[17] = 0.0
    src = torch.from_numpy(x).to(DEV)
    dst = torch.empty(x.shape, dtype=torch.uint8, device=DEV)
    scale = cg.quantize(src, dst, cg.FP8_E4M3)
    amax = float(np.max(np.abs(x)))
    assert scale == np.float32(amax) / np.float32(448.0)
    inv = np.float32(1.0) / np.float32(scale)
    ref = torch.from_numpy(np.clip(x * inv, -448, 448).astype(np.float32)).to(torch.float8_e4m3fn).view(torch.uint8)
    np.testing.assert_array_equal(dst.cpu().numpy(), ref.numpy())
    # all-zero tensor: scale 1
    z = torch.zeros(1000, device=DEV)
    zd = torch.empty(1000, dtype=torch.uint8, device=DEV)
    assert cg.quantize(z, zd, cg.FP8_E4M3) == 1.0
    assert not zd.cpu().numpy().any()


SHAPES = [
    Shape(2, 12, 11, 32, 48, 3, 3, 1, 1, 1, 1, 1, 1),
    Shape(2, 9, 9, 24, 40, 3, 3, 1, 0, 2, 2, 1, 1),
    Shape(1, 14, 14, 64, 64, 1, 1, 0, 0, 1, 1, 1, 1),
    Shape(2, 10, 10, 5, 12, 3, 3, 1, 1, 1, 1, 2, 2),   # channels not a multiple of the chunk
]


def _run(desc, in_d, x_d, out_shape, dtype, scales=None, out0=None, alpha=1.0, beta=0.0, epi=None):
    out = torch.zeros(out_shape, device=DEV) if out0 is None else torch.from_numpy(out0.astype(np.float32)).to(DEV)
    e = {} if epi is None else {k: torch.from_numpy(np.asarray(v, np.float32)).to(DEV) for k, v in epi.items()
                                 if k != "relu"}
    cg.conv_ex(desc, in_d, x_d, out, alpha=alpha, beta=beta, dtype=dtype, scales=scales,
               relu=bool(epi and epi.get("relu")), **e)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("layout", [NHWC, NCHW])
@pytest.mark.parametrize("si", range(len(SHAPES)))
def test_bf16_all_operations(layout, si):
    s = SHAPES[si]
    rng = np.random.default_rng(300 + si)
    X = rng.standard_normal(s.x_shape(layout)).astype(np.float32)
    W = (rng.standard_normal(s.w_shape(layout)) * np.sqrt(2.0 / s.K)).astype(np.float32)
    dY = rng.standard_normal(s.y_shape(layout)).astype(np.float32)
    Xd, Xq = q_bf16(X)
    Wd, Wq = q_bf16(W)
    Yd, Yq = q_bf16(dY)
    epi = dict(bias=0.1 * rng.standard_normal(s.kn), relu=True)
    geo = (s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil, s.hdil)
    print(f"bf16 layout={layout} {s}", flush=True)
    y = _run(cg.make_desc(cg.FWD, layout, *geo), Wd, Xd, s.y_shape(layout), cg.BF16, epi=epi)
    assert rel_err(y, oracle.fwd(layout, s, Wq, Xq, **epi))[0] <= 2e-5
    assert rel_err(y, oracle.fwd(layout, s, W, X, **epi))[0] <= 2e-2
    dw = _run(cg.make_desc(cg.WGRAD, layout, *geo), Yd, Xd, s.w_shape(layout), cg.BF16)
    assert rel_err(dw, oracle.wgrad(layout, s, Yq, Xq))[0] <= 2e-5
    dX0 = rng.standard_normal(s.x_shape(layout))
    dx = _run(cg.make_desc(cg.DGRAD, layout, *geo), Wd, Yd, s.x_shape(layout), cg.BF16, out0=dX0, beta=1.0)
    assert rel_err(dx, oracle.dgrad(layout, s, Wq, Yq, dX0=dX0.astype(np.float32)))[0] <= 2e-5


@pytest.mark.parametrize("layout", [NHWC, NCHW])
@pytest.mark.parametrize("si", range(len(SHAPES)))
def test_fp8_forward(layout, si):
    s = SHAPES[si]
    rng = np.random.default_rng(400 + si)
    X = np.abs(rng.standard_normal(s.x_shape(layout))).astype(np.float32)   # post-ReLU activations
    W = (rng.standard_normal(s.w_shape(layout)) * 0.02).astype(np.float32)
    Xd, Xq, sx = q_fp8(X)
    Wd, Wq, sw = q_fp8(W)
    geo = (s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil, s.hdil)
    print(f"fp8 layout={layout} {s}", flush=True)
    y = _run(cg.make_desc(cg.FWD, layout, *geo), Wd, Xd, s.y_shape(layout), cg.FP8_E4M3, scales=(sw, sx))
    assert rel_err(y, oracle.fwd(layout, s, Wq, Xq))[0] <= FP8_TOL
    assert rel_err(y, oracle.fwd(layout, s, W, X))[0] <= 1.5e-1


def test_fp8_rejects_backward_and_missing_scales():
    s = SHAPES[0]
    geo = (s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil, s.hdil)
    buf = torch.zeros(4096, device=DEV)
    with pytest.raises(cg.ConvGemmError) as e:
        cg.conv_ex(cg.make_desc(cg.DGRAD, NHWC, *geo), buf, buf, buf, dtype=cg.FP8_E4M3, scales=(1.0, 1.0))
    assert e.value.status == 5
    with pytest.raises(cg.ConvGemmError) as e:
        cg.conv_ex(cg.make_desc(cg.FWD, NHWC, *geo), buf, buf, buf, dtype=cg.FP8_E4M3)
    assert e.value.status == 1


@pytest.mark.parametrize("dtype", ["bf16", "fp8"])
def test_every_lowp_configuration(dtype):
    elem = cg.BF16 if dtype == "bf16" else cg.FP8_E4M3
    ncfg = cg.num_tile_shapes(elem)
    s = Shape(2, 11, 10, 64, 80, 3, 3, 1, 1, 1, 1, 1, 1)
    rng = np.random.default_rng(5)
    for layout in (NHWC, NCHW):
        X = rng.standard_normal(s.x_shape(layout)).astype(np.float32)
        W = (rng.standard_normal(s.w_shape(layout)) * 0.1).astype(np.float32)
        if dtype == "bf16":
            (Xd, Xq), (Wd, Wq), scales = q_bf16(X), q_bf16(W), None
        else:
            Xd, Xq, sx = q_fp8(X)
            Wd, Wq, sw = q_fp8(W)
            scales = (sw, sx)
        ref = oracle.fwd(layout, s, Wq, Xq)
        geo = (s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil, s.hdil)
        for cfg in range(ncfg):
            for splits in (1, 3):
                cg.force_config(cfg, splits)
                try:
                    y = _run(cg.make_desc(cg.FWD, layout, *geo), Wd, Xd, s.y_shape(layout), elem, scales=scales)
                finally:
                    cg.force_config(-1, 0)
                assert rel_err(y, ref)[0] <= (2e-5 if dtype == "bf16" else FP8_TOL), (layout, cfg, splits)


@pytest.mark.parametrize("splits", [1, 3, 0])
def test_every_bf16_configuration_backward(splits):
    """Weight gradient and backward data through every BF16 configuration that
    offers them: 16-byte channel groups (C, kn multiples of 8; the weight
    gradient's transposing loader in NHWC) and element loads (C = 12), with
    split-K and stream-K (splits 0, weight gradient)."""
    ncfg = cg.num_tile_shapes(cg.BF16)
    cases = [(Shape(2, 11, 10, 64, 80, 3, 3, 1, 1, 1, 1, 1, 1), True),
             (Shape(2, 9, 8, 12, 24, 3, 3, 1, 1, 2, 1, 1, 1), False),
             (Shape(3, 7, 7, 40, 16, 1, 1, 0, 0, 1, 1, 1, 1), True)]
    for ci, (s, vec) in enumerate(cases):
        rng = np.random.default_rng(700 + ci)
        geo = (s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil, s.hdil)
        for layout in (NHWC, NCHW):
            X = rng.standard_normal(s.x_shape(layout)).astype(np.float32)
            W = (rng.standard_normal(s.w_shape(layout)) * 0.1).astype(np.float32)
            dY = rng.standard_normal(s.y_shape(layout)).astype(np.float32)
            (Xd, Xq), (Wd, Wq), (Yd, Yq) = q_bf16(X), q_bf16(W), q_bf16(dY)
            ref_w = oracle.wgrad(layout, s, Yq, Xq)
            ref_x = oracle.dgrad(layout, s, Wq, Yq)
            for cfg in range(ncfg):
                for op in (cg.WGRAD, cg.DGRAD):
                    if not cg.tile_supports(cg.BF16, cfg, op, layout, vec):
                        continue
                    print(f"bf16 backward cfg={cfg} splits={splits} op={op} layout={layout} {s}", flush=True)
                    cg.force_config(cfg, splits)
                    try:
                        if op == cg.WGRAD:
                            got = _run(cg.make_desc(op, layout, *geo), Yd, Xd, s.w_shape(layout), cg.BF16)
                            ref = ref_w
                        else:
                            got = _run(cg.make_desc(op, layout, *geo), Wd, Yd, s.x_shape(layout), cg.BF16)
                            ref = ref_x
                        assert cg.last_plan()[0] == cfg
                    finally:
                        cg.force_config(-1, 0)
                    assert rel_err(got, ref)[0] <= 2e-5, (cfg, splits, op, layout, ci)


# Shapes for the pipelined forward kernel (128-byte K steps, 2-4 LDS stages):
# a single K step, every tail length, ragged channel blocks, padding, stride,
# dilation, M and N tails.
LP_SHAPES = [
    Shape(2, 9, 9, 64, 40, 1, 1, 0, 0, 1, 1, 1, 1),      # bf16: one K step
    Shape(2, 11, 10, 72, 80, 3, 3, 1, 1, 1, 1, 1, 1),    # ragged channel block (144 B / 72 B)
    Shape(1, 13, 12, 128, 96, 3, 3, 1, 1, 2, 2, 1, 1),   # stride 2
    Shape(2, 10, 10, 48, 136, 3, 3, 2, 2, 1, 1, 2, 2),   # dilation, N tail
    Shape(1, 7, 7, 256, 64, 5, 5, 2, 2, 1, 1, 1, 1),     # 25 taps x 4/2 blocks
    Shape(3, 6, 5, 32, 16, 1, 1, 0, 0, 1, 1, 1, 1),      # tiny, M tail
]


@pytest.mark.parametrize("dtype", ["f32", "bf16", "fp8"])
@pytest.mark.parametrize("si", range(len(LP_SHAPES)))
def test_pipelined_forward_configurations(dtype, si):
    elem = {"f32": cg.F32, "bf16": cg.BF16, "fp8": cg.FP8_E4M3}[dtype]
    lp = [i for i, t in enumerate(cg.tile_configs(elem)) if t[5] > 0 and t[6] == 0 and t[7] == 0 and cg.tile_supports(elem, i, cg.FWD, NHWC)]
    assert lp, "no pipelined configurations"
    s = LP_SHAPES[si]
    if dtype == "fp8" and s.c % 16:
        pytest.skip("FP8 16-byte chunks need C % 16 == 0")
    rng = np.random.default_rng(50 + si)
    X = rng.standard_normal(s.x_shape(NHWC)).astype(np.float32)
    W = (rng.standard_normal(s.w_shape(NHWC)) * 0.1).astype(np.float32)
    if dtype == "f32":
        Xd, Xq, Wd, Wq, scales = torch.from_numpy(X).to(DEV), X, torch.from_numpy(W).to(DEV), W, None
    elif dtype == "bf16":
        (Xd, Xq), (Wd, Wq), scales = q_bf16(X), q_bf16(W), None
    else:
        Xd, Xq, sx = q_fp8(X)
        Wd, Wq, sw = q_fp8(W)
        scales = (sw, sx)
    tol = {"f32": 1e-5, "bf16": 2e-5, "fp8": FP8_TOL}[dtype]
    epi = dict(bias=(0.1 * rng.standard_normal(s.kn)).astype(np.float32), relu=True)
    ref = oracle.fwd(NHWC, s, Wq, Xq, **epi)
    geo = (s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil, s.hdil)
    d = cg.make_desc(cg.FWD, NHWC, *geo)
    bias = torch.from_numpy(epi["bias"]).to(DEV)
    for cfg in lp:
        for splits in (1, 2, 3):
            cg.force_config(cfg, splits)
            try:
                Y = torch.full(s.y_shape(NHWC), float("nan"), device=DEV)
                print(f"lp {dtype} shape {si} cfg {cfg} splits {splits}", flush=True)
                cg.conv_ex(d, Wd, Xd, Y, dtype=elem, scales=scales, bias=bias, relu=True)
                torch.cuda.synchronize()
                assert cg.last_plan()[0] == cfg
            finally:
                cg.force_config(-1, 0)
            y = Y.cpu().numpy()
            assert rel_err(y, ref)[0] <= tol, (cfg, splits)


@pytest.mark.parametrize("si", range(len(LP_SHAPES)))
def test_pipelined_forward_nchw_fp32(si):
    """The pipelined NCHW forward kernel (FP32, 4-byte activation loads into a
    K-major LDS tile) on every tail, padding, stride and dilation case, with
    the fused epilogue and split-K, against the FP64 oracle."""
    lpn = [i for i, t in enumerate(cg.tile_configs(cg.F32))
           if t[5] > 0 and t[6] == 0 and cg.tile_supports(cg.F32, i, cg.FWD, NCHW)]
    assert lpn, "no pipelined NCHW configurations"
    s = LP_SHAPES[si]
    rng = np.random.default_rng(70 + si)
    X = rng.standard_normal(s.x_shape(NCHW)).astype(np.float32)
    W = (rng.standard_normal(s.w_shape(NCHW)) * 0.1).astype(np.float32)
    epi = dict(bias=(0.1 * rng.standard_normal(s.kn)).astype(np.float32), relu=True)
    ref = oracle.fwd(NCHW, s, W, X, **epi)
    geo = (s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil, s.hdil)
    d = cg.make_desc(cg.FWD, NCHW, *geo)
    Xd, Wd = torch.from_numpy(X).to(DEV), torch.from_numpy(W).to(DEV)
    bias = torch.from_numpy(epi["bias"]).to(DEV)
    for cfg in lpn:
        for splits in (1, 2, 3):
            cg.force_config(cfg, splits)
            try:
                Y = torch.full(s.y_shape(NCHW), float("nan"), device=DEV)
                print(f"lp nchw f32 shape {si} cfg {cfg} splits {splits}", flush=True)
                cg.conv_ex(d, Wd, Xd, Y, bias=bias, relu=True)
                torch.cuda.synchronize()
                assert cg.last_plan()[0] == cfg
            finally:
                cg.force_config(-1, 0)
            assert rel_err(Y.cpu().numpy(), ref)[0] <= 1e-5, (cfg, splits)


# Stream-K shapes: several tiles per workgroup range, fewer units than
# workgroups, ragged N tiles, long K per tile.
SK_LP_SHAPES = [
    Shape(4, 14, 14, 64, 64, 3, 3, 1, 1, 1, 1, 1, 1),
    Shape(2, 9, 7, 32, 20, 3, 3, 1, 1, 2, 1, 1, 1),
    Shape(8, 28, 28, 32, 96, 1, 1, 0, 0, 1, 1, 1, 1),
    Shape(3, 16, 16, 48, 24, 5, 5, 2, 2, 1, 1, 1, 1),
]


@pytest.mark.parametrize("dtype,layout", [("f32", NHWC), ("f32", NCHW), ("bf16", NHWC), ("fp8", NHWC)])
def test_pipelined_stream_k(dtype, layout):
    """Stream-K (split-K 0) on every pipelined configuration: pieces of tiles
    shared between workgroups are summed by the fix-up kernel; forward with
    bias + ReLU against the FP64 oracle."""
    elem = {"f32": cg.F32, "bf16": cg.BF16, "fp8": cg.FP8_E4M3}[dtype]
    cfgs = [i for i, t in enumerate(cg.tile_configs(elem)) if t[5] > 0 and t[6] == 0 and t[7] == 0 and cg.tile_supports(elem, i, cg.FWD, layout)]
    assert cfgs
    tol = {"f32": 1e-5, "bf16": 2e-5, "fp8": FP8_TOL}[dtype]
    for si, s in enumerate(SK_LP_SHAPES):
        if s.c % {"f32": 4, "bf16": 8, "fp8": 16}[dtype]:
            continue
        rng = np.random.default_rng(1300 + si)
        X = rng.standard_normal(s.x_shape(layout)).astype(np.float32)
        W = (rng.standard_normal(s.w_shape(layout)) * 0.1).astype(np.float32)
        if dtype == "f32":
            Xd, Xq, Wd, Wq, scales = torch.from_numpy(X).to(DEV), X, torch.from_numpy(W).to(DEV), W, None
        elif dtype == "bf16":
            (Xd, Xq), (Wd, Wq), scales = q_bf16(X), q_bf16(W), None
        else:
            Xd, Xq, sx = q_fp8(X)
            Wd, Wq, sw = q_fp8(W)
            scales = (sw, sx)
        bias = (0.1 * rng.standard_normal(s.kn)).astype(np.float32)
        ref = oracle.fwd(layout, s, Wq, Xq, bias=bias, relu=True)
        geo = (s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil, s.hdil)
        d = cg.make_desc(cg.FWD, layout, *geo)
        bias_d = torch.from_numpy(bias).to(DEV)
        for cfg in cfgs:
            cg.force_config(cfg, 0)
            try:
                Y = torch.full(s.y_shape(layout), float("nan"), device=DEV)
                print(f"stream-K {dtype} layout {layout} shape {si} cfg {cfg}", flush=True)
                cg.conv_ex(d, Wd, Xd, Y, dtype=elem, scales=scales, bias=bias_d, relu=True)
                torch.cuda.synchronize()
                plan = cg.last_plan()
                # FP8 wave tiles of 128x64 and up have no stream-K form: plain tiles
                t = cg.tile_configs(elem)[cfg]
                no_sk = dtype == "fp8" and (t[0] // t[3]) * (t[1] // t[4]) >= 128 * 64
                assert plan[0] == cfg and plan[1] == (1 if no_sk else 0), plan
            finally:
                cg.force_config(-1, 0)
            assert rel_err(Y.cpu().numpy(), ref)[0] <= tol, (si, cfg)
