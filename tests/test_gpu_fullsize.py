"""Full-size checks on the ResNet-50 layer sweep at batch 32 (L1-L5).

At these sizes the direct FP64 oracle is too slow, so the checks are:
* the FP64 oracle at 256 sampled output elements per operation
  (oracle/sampled.py, pinned against the direct oracle);
* the adjoint identity <conv(X,W), dY> = <X, dgrad(W,dY)> = <W, wgrad(X,dY)>
  evaluated in float64 from the FP32 GPU results.
"""
import numpy as np
import pytest
import torch

from oracle import NCHW, NHWC, Shape, sampled

pytestmark = pytest.mark.gpu

import convgemm_amd as cg  # noqa: E402

LAYERS = {
    "L1_1x1_256to64_56": Shape(32, 56, 56, 256, 64, 1, 1, 0, 0, 1, 1, 1, 1),
    "L2_3x3_64to64_56": Shape(32, 56, 56, 64, 64, 3, 3, 1, 1, 1, 1, 1, 1),
    "L3_3x3s2_128to128_56": Shape(32, 56, 56, 128, 128, 3, 3, 1, 1, 2, 2, 1, 1),
    "L4_1x1_1024to256_14": Shape(32, 14, 14, 1024, 256, 1, 1, 0, 0, 1, 1, 1, 1),
    "L5_3x3_512to512_7": Shape(32, 7, 7, 512, 512, 3, 3, 1, 1, 1, 1, 1, 1),
}
TOL = 1e-5


def _dims(s, layout):
    return (s.b, s.h, s.w, s.c) if layout == NHWC else (s.b, s.c, s.h, s.w)


@pytest.mark.parametrize("layout", [NHWC, NCHW])
@pytest.mark.parametrize("name", list(LAYERS))
def test_fullsize_layer(name, layout):
    s = LAYERS[name]
    print(f"full-size {name} layout={layout}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(sum(map(ord, name)) * 2 + layout)
    X = torch.randn(s.x_shape(layout), device="cuda", generator=g)
    W = torch.randn(s.w_shape(layout), device="cuda", generator=g) * (2.0 / s.K) ** 0.5
    dY = torch.randn(s.y_shape(layout), device="cuda", generator=g)
    kn = s.kn
    bias = 0.1 * torch.randn(kn, device="cuda", generator=g)
    mean = 0.1 * torch.randn(kn, device="cuda", generator=g)
    inv_std = 1.0 / torch.sqrt(0.5 + 1.5 * torch.rand(kn, device="cuda", generator=g) + 1e-5)
    gamma = 0.5 + torch.rand(kn, device="cuda", generator=g)
    bnb = 0.1 * torch.randn(kn, device="cuda", generator=g)

    fwd = cg.sconvGemmNHWC if layout == NHWC else cg.sconvGemmNCHW
    back = cg.sconvGemmNHWC_back if layout == NHWC else cg.sconvGemmNCHW_back
    geo = (s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil, s.hdil)
    Ye = torch.empty(s.y_shape(layout), device="cuda")
    fwd("N", *_dims(s, layout), *geo, W, X, Ye, bias, mean, inv_std, gamma, bnb, True)
    Y = torch.empty(s.y_shape(layout), device="cuda")
    fwd("N", *_dims(s, layout), *geo, W, X, Y)
    dW = torch.empty(s.w_shape(layout), device="cuda")
    fwd("T", *_dims(s, layout), *geo, dY, X, dW)
    dX = torch.zeros(s.x_shape(layout), device="cuda")
    back(*_dims(s, layout), s.kn, s.kh, s.kw, s.vstr, s.hstr, s.vpad, s.hpad, s.vdil, s.hdil, W, dY, dX)
    torch.cuda.synchronize()

    # adjoint identity in float64
    a = torch.vdot(Y.double().flatten(), dY.double().flatten()).item()
    b = torch.vdot(X.double().flatten(), dX.double().flatten()).item()
    c = torch.vdot(W.double().flatten(), dW.double().flatten()).item()
    scale = (Y.double().norm() * dY.double().norm()).item()
    assert abs(a - b) <= 1e-5 * scale and abs(a - c) <= 1e-5 * scale, (a, b, c, scale)

    rng = np.random.default_rng(11)
    Xh, Wh, dYh = X.cpu().numpy(), W.cpu().numpy(), dY.cpu().numpy()
    epi = dict(bias=bias.cpu().numpy(), mean=mean.cpu().numpy(), inv_std=inv_std.cpu().numpy(),
               gamma=gamma.cpu().numpy(), bnbeta=bnb.cpu().numpy(), relu=True)
    ys = np.stack([rng.integers(0, n, 256) for n in (s.b, s.ho, s.wo, s.kn)], 1)
    ref = sampled.fwd_sampled(layout, s, Wh, Xh, ys, **epi)
    got = sampled.nhwc_index(layout, Ye.cpu().numpy(), ys)
    ref_noepi = sampled.fwd_sampled(layout, s, Wh, Xh, ys)
    got_noepi = sampled.nhwc_index(layout, Y.cpu().numpy(), ys)
    assert np.max(np.abs(got - ref)) <= TOL * max(np.max(np.abs(ref)), 1.0)
    assert np.max(np.abs(got_noepi - ref_noepi)) <= TOL * np.max(np.abs(ref_noepi))
    ws = np.stack([rng.integers(0, n, 48) for n in (s.c, s.kh, s.kw, s.kn)], 1)
    ref = sampled.wgrad_sampled(layout, s, dYh, Xh, ws)
    got = sampled.w_index(layout, dW.cpu().numpy(), ws)
    assert np.max(np.abs(got - ref)) <= TOL * np.max(np.abs(ref))
    xs = np.stack([rng.integers(0, n, 128) for n in (s.b, s.h, s.w, s.c)], 1)
    ref = sampled.dgrad_sampled(layout, s, Wh, dYh, xs)
    got = sampled.nhwc_index(layout, dX.cpu().numpy(), xs)
    assert np.max(np.abs(got - ref)) <= TOL * np.max(np.abs(ref))


def _gather(layout, Y, ys):
    """Y (on the GPU) at NHWC-ordered (b, x, y, k) samples, copied back."""
    i = [torch.as_tensor(ys[:, j], device=Y.device) for j in range(4)]
    v = Y[i[0], i[1], i[2], i[3]] if layout == NHWC else Y[i[0], i[3], i[1], i[2]]
    return v.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("layout", [NHWC, NCHW])
@pytest.mark.parametrize("name", list(LAYERS))
def test_fullsize_pipelined_forward(name, layout):
    """Every pipelined FP32 forward configuration of this layout, forced on the
    full-size layer with the fused epilogue (split-K 1 and 3, and stream-K,
    which on these tile counts runs data-parallel tiles and a stream-K
    remainder): 256 sampled outputs against the FP64 oracle."""
    s = LAYERS[name]
    cfgs = [i for i, t in enumerate(cg.tile_configs(cg.F32)[:cg.num_tile_shapes(cg.F32)])
            if t[5] > 0 and t[7] == 0 and cg.tile_supports(cg.F32, i, cg.FWD, layout)]
    assert cfgs
    g = torch.Generator(device="cuda").manual_seed(sum(map(ord, name)) * 3 + layout)
    X = torch.randn(s.x_shape(layout), device="cuda", generator=g)
    W = torch.randn(s.w_shape(layout), device="cuda", generator=g) * (2.0 / s.K) ** 0.5
    kn = s.kn
    bias = 0.1 * torch.randn(kn, device="cuda", generator=g)
    mean = 0.1 * torch.randn(kn, device="cuda", generator=g)
    inv_std = 1.0 / torch.sqrt(0.5 + 1.5 * torch.rand(kn, device="cuda", generator=g) + 1e-5)
    gamma = 0.5 + torch.rand(kn, device="cuda", generator=g)
    bnb = 0.1 * torch.randn(kn, device="cuda", generator=g)
    rng = np.random.default_rng(12)
    ys = np.stack([rng.integers(0, n, 256) for n in (s.b, s.ho, s.wo, s.kn)], 1)
    epi = dict(bias=bias.cpu().numpy(), mean=mean.cpu().numpy(), inv_std=inv_std.cpu().numpy(),
               gamma=gamma.cpu().numpy(), bnbeta=bnb.cpu().numpy(), relu=True)
    ref = sampled.fwd_sampled(layout, s, W.cpu().numpy(), X.cpu().numpy(), ys, **epi)
    d = cg.make_desc(cg.FWD, layout, s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr,
                     s.vdil, s.hdil)
    Y = torch.empty(s.y_shape(layout), device="cuda")
    for cfg in cfgs:
        for splits in (1, 3, 0):  # 0: stream-K (data-parallel tiles + a stream-K remainder)
            print(f"full-size pipelined {name} layout={layout} cfg {cfg} splits {splits}", flush=True)
            Y.fill_(float("nan"))
            cg.force_config(cfg, splits)
            try:
                cg.conv_ex(d, W, X, Y, bias=bias, mean=mean, inv_std=inv_std, gamma=gamma, bnbeta=bnb, relu=True)
                torch.cuda.synchronize()
                assert cg.last_plan()[0] == cfg
            finally:
                cg.force_config(-1, 0)
            got = _gather(layout, Y, ys)
            assert np.max(np.abs(got - ref)) <= TOL * max(np.max(np.abs(ref)), 1.0), (cfg, splits)


@pytest.mark.parametrize("layout", [NHWC, NCHW])
@pytest.mark.parametrize("name", ["L1_1x1_256to64_56", "L3_3x3s2_128to128_56", "L4_1x1_1024to256_14"])
def test_fullsize_stream_k_general_kernel(name, layout):
    """Stream-K on the general kernel (forward with the fused epilogue, and the
    weight gradient), forced on full-size layers whose 64x64 tile counts are
    not multiples of the 256 CUs, so data-parallel tiles, stream-K pieces and
    the fix-up all run: sampled outputs against the FP64 oracle."""
    s = LAYERS[name]
    cfg = next(i for i, t in enumerate(cg.tile_configs(cg.F32)) if t[:2] == (64, 64) and t[2] == 32 and t[5] == 0)
    g = torch.Generator(device="cuda").manual_seed(sum(map(ord, name)) * 5 + layout)
    X = torch.randn(s.x_shape(layout), device="cuda", generator=g)
    W = torch.randn(s.w_shape(layout), device="cuda", generator=g) * (2.0 / s.K) ** 0.5
    dY = torch.randn(s.y_shape(layout), device="cuda", generator=g)
    bias = 0.1 * torch.randn(s.kn, device="cuda", generator=g)
    geo = (s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil, s.hdil)
    Y = torch.full(s.y_shape(layout), float("nan"), device="cuda")
    dW = torch.full(s.w_shape(layout), float("nan"), device="cuda")
    print(f"stream-K general kernel {name} layout={layout} cfg {cfg}", flush=True)
    cg.force_config(cfg, 0)
    try:
        cg.conv_ex(cg.make_desc(cg.FWD, layout, *geo), W, X, Y, bias=bias, relu=True)
        torch.cuda.synchronize()
        plan = cg.last_plan()  # (a weight repack), the stream-K kernel, the fix-up
        assert plan[:2] == (cfg, 0) and plan[2] >= 2, plan
        cg.conv_ex(cg.make_desc(cg.WGRAD, layout, *geo), dY, X, dW)
        torch.cuda.synchronize()
        assert cg.last_plan()[:2] == (cfg, 0), cg.last_plan()
    finally:
        cg.force_config(-1, 0)
    rng = np.random.default_rng(13)
    Xh = X.cpu().numpy()
    ys = np.stack([rng.integers(0, n, 256) for n in (s.b, s.ho, s.wo, s.kn)], 1)
    ref = sampled.fwd_sampled(layout, s, W.cpu().numpy(), Xh, ys, bias=bias.cpu().numpy(), relu=True)
    got = _gather(layout, Y, ys)
    assert np.max(np.abs(got - ref)) <= TOL * max(np.max(np.abs(ref)), 1.0)
    ws = np.stack([rng.integers(0, n, 48) for n in (s.c, s.kh, s.kw, s.kn)], 1)
    ref = sampled.wgrad_sampled(layout, s, dY.cpu().numpy(), Xh, ws)
    got = sampled.w_index(layout, dW.cpu().numpy(), ws)
    assert np.max(np.abs(got - ref)) <= TOL * np.max(np.abs(ref))
