"""GPU parity of the FP32 implicit-GEMM path against the oracle.

Every test calls the HIP kernels through the C ABI (libconvgemm.so) and
compares with the FP64 direct-convolution oracle (pinned in test_oracle.py).
FP32 bar: max|y - y64| / max|y64| <= 1e-5 (normwise-max; the reference's own
element-wise check is used on its own U[0,1] data in
test_reference_acceptance_uniform_data).

Edge cases follow what the reference's tests and its code call for: every
epilogue combination including NULL pointers, alpha/beta, dilation,
asymmetric stride and padding, multi-block K, channel counts that break
16-byte vectors, kernels larger than the stride, dX accumulation, empty
problems, and every tile configuration with and without split-K.
"""
import numpy as np
import pytest
import torch

import oracle
from oracle import NCHW, NHWC, Shape

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    import convgemm_amd as cg
    cg.lib()  # fail loudly if the HIP library is missing
    yield
    cg.force_config(-1, 0)


from tests.gpu_util import (DEV, dev, host, random_problem, rel_err, run_dgrad, run_fwd,  # noqa: E402
                            run_wgrad)
import convgemm_amd as cg  # noqa: E402

SHAPES = [
    Shape(1, 56, 56, 64, 64, 3, 3, 1, 1, 1, 1, 1, 1),   # the reference's CPU configuration (batch 1)
    Shape(2, 9, 7, 5, 6, 3, 3, 1, 1, 1, 1, 1, 1),       # C, kn not multiples of 4
    Shape(2, 10, 11, 8, 12, 3, 3, 2, 1, 1, 2, 2, 1),    # dilation, asymmetric stride and padding
    Shape(3, 8, 8, 16, 20, 1, 1, 0, 0, 1, 1, 1, 1),     # 1x1
    Shape(2, 15, 13, 12, 16, 5, 5, 2, 2, 2, 2, 1, 1),   # 5x5 s2
    Shape(1, 23, 23, 3, 16, 11, 11, 0, 0, 4, 4, 1, 1),  # AlexNet-like first layer, C = 3
    Shape(2, 14, 14, 40, 24, 3, 3, 1, 1, 3, 3, 1, 1),   # stride 3 > half kernel
    Shape(1, 6, 6, 4, 4, 6, 6, 0, 0, 1, 1, 1, 1),       # full window
    Shape(2, 7, 7, 100, 36, 3, 3, 1, 1, 1, 1, 1, 1),    # K = 900, several K tiles, ragged
    Shape(1, 12, 9, 8, 8, 2, 3, 0, 1, 2, 3, 3, 2),      # even kernel, stride 3, dilation 3
    Shape(1, 1, 1, 7, 5, 1, 1, 0, 0, 1, 1, 1, 1),       # single pixel
    Shape(2, 16, 16, 32, 32, 7, 7, 3, 3, 2, 2, 1, 1),   # 7x7 s2 p3
]


def _case(rng, s, layout):
    print(f"case layout={layout} {s}", flush=True)
    return random_problem(rng, s, layout)


@pytest.mark.parametrize("layout", [NHWC, NCHW])
@pytest.mark.parametrize("si", range(len(SHAPES)))
def test_forward_epilogue(layout, si):
    s = SHAPES[si]
    rng = np.random.default_rng(1000 + si)
    X, W, _, epi = _case(rng, s, layout)
    Y = run_fwd(layout, s, W, X, epi)
    ref = oracle.fwd(layout, s, W, X, **epi)
    mx, nrm = rel_err(Y, ref)
    assert mx <= TOL, (mx, nrm)


@pytest.mark.parametrize("layout", [NHWC, NCHW])
@pytest.mark.parametrize("si", range(len(SHAPES)))
def test_weight_gradient(layout, si):
    s = SHAPES[si]
    rng = np.random.default_rng(2000 + si)
    X, _, dY, _ = _case(rng, s, layout)
    dW = run_wgrad(layout, s, dY, X)
    mx, nrm = rel_err(dW, oracle.wgrad(layout, s, dY, X))
    assert mx <= TOL, (mx, nrm)


@pytest.mark.parametrize("layout", [NHWC, NCHW])
@pytest.mark.parametrize("si", range(len(SHAPES)))
def test_backward_data_accumulates(layout, si):
    s = SHAPES[si]
    rng = np.random.default_rng(3000 + si)
    _, W, dY, _ = _case(rng, s, layout)
    dX0 = rng.standard_normal(s.x_shape(layout)).astype(np.float32)
    dX = run_dgrad(layout, s, W, dY, dX0)
    ref = oracle.dgrad(layout, s, W, dY, dX0=dX0)  # dx += row2im(W^T dY)
    mx, nrm = rel_err(dX, ref)
    assert mx <= TOL, (mx, nrm)


CFG_SHAPES = [Shape(2, 11, 10, 24, 40, 3, 3, 1, 1, 1, 1, 1, 1), Shape(1, 13, 12, 20, 36, 3, 3, 1, 0, 2, 2, 1, 1)]


# every configuration of the general kernel (the pipelined ones have tests of their own)
GENERAL_CFGS = [c for c, t in enumerate(cg.tile_configs(cg.F32)) if t[5] == 0 and t[6] == 0]


@pytest.mark.parametrize("splits", [0, 1, 3])
@pytest.mark.parametrize("cfg", GENERAL_CFGS)
def test_every_tile_configuration(cfg, splits):
    """Force each kernel instantiation (and the split-K reduction; splits 0 =
    stream-K with its fix-up, forward and weight gradient) on every op.
    A configuration an operation does not offer falls back to the planner."""
    cg.force_config(cfg, splits)
    try:
        for si, s in enumerate(CFG_SHAPES):
            for layout in (NHWC, NCHW):
                rng = np.random.default_rng(cfg * 100 + splits * 10 + si)
                X, W, dY, epi = _case(rng, s, layout)
                mx, _ = rel_err(run_fwd(layout, s, W, X, epi), oracle.fwd(layout, s, W, X, **epi))
                assert mx <= TOL, ("fwd", layout, si, mx)
                mx, _ = rel_err(run_wgrad(layout, s, dY, X), oracle.wgrad(layout, s, dY, X))
                assert mx <= TOL, ("wgrad", layout, si, mx)
                dX0 = np.zeros(s.x_shape(layout), np.float32)
                mx, _ = rel_err(run_dgrad(layout, s, W, dY, dX0), oracle.dgrad(layout, s, W, dY))
                assert mx <= TOL, ("dgrad", layout, si, mx)
    finally:
        cg.force_config(-1, 0)


# a shape with more than 8 M tiles and several N tiles for every operation (64x64 tiles)
RASTER_SHAPE = Shape(2, 18, 18, 96, 160, 3, 3, 1, 1, 1, 1, 1, 1)


@pytest.mark.parametrize("layout", [NHWC, NCHW])
def test_raster_orders(layout):
    """Raster orders 1 (M panels adjacent) and 2 (bands of 8 M panels) only change
    which workgroup computes which tile: every operation, on general and
    pipelined configurations, with split-K 1 and 3 is bit-identical to order 0
    (a tile's arithmetic does not depend on its position) and within 1e-5 of the
    oracle.  Under stream-K a tile's K range is cut where its position in the
    order puts it, so there every order is checked against the oracle only."""
    s = RASTER_SHAPE
    rng = np.random.default_rng(321 + layout)
    X, W, dY, epi = _case(rng, s, layout)
    refs = {"fwd": oracle.fwd(layout, s, W, X, **epi), "wgrad": oracle.wgrad(layout, s, dY, X),
            "dgrad": oracle.dgrad(layout, s, W, dY)}
    confs = cg.tile_configs(cg.F32)
    ns = cg.num_tile_shapes(cg.F32)
    opk = {"fwd": 0 + layout, "wgrad": 2 + layout, "dgrad": 4 + layout}
    shapes = [b for b in range(ns) if confs[b][:2] == (64, 64) and confs[b][5] == 0][:1]
    shapes += [b for b in range(ns) if confs[b][5] > 0 and confs[b][7] == 0 and cg.tile_supports(cg.F32, b, cg.FWD, layout)][:1]
    assert len(shapes) == 2
    dX0 = np.zeros(s.x_shape(layout), np.float32)
    for base in shapes:
        for splits in (1, 3, 0):
            for op, run in (("fwd", lambda: run_fwd(layout, s, W, X, epi)), ("wgrad", lambda: run_wgrad(layout, s, dY, X)),
                            ("dgrad", lambda: run_dgrad(layout, s, W, dY, dX0))):
                if not cg.lib().cg_tilecfg_supports(cg.F32, base, opk[op], 1):
                    continue
                outs = []
                for r in range(3):
                    cfg = base + r * ns
                    print(f"raster {r} cfg {cfg} {op} layout={layout} splits={splits}", flush=True)
                    cg.force_config(cfg, splits)
                    try:
                        outs.append(run())
                        assert cg.last_plan()[0] == cfg
                    finally:
                        cg.force_config(-1, 0)
                for r in range(3):
                    assert rel_err(outs[r], refs[op])[0] <= TOL, (op, base, splits, r)
                    if splits:
                        assert np.array_equal(outs[r], outs[0]), (op, base, splits, r)


@pytest.mark.parametrize("splits", [8, 16, 40])
def test_deep_split_k_reduction(splits):
    """Split-K with 8 or more slices reduces each output in four interleaved groups
    (combined in a fixed order): weight gradient in both layouts, forward with the
    epilogue, strided backward data, against the oracle; and a one-problem batch
    (the grouped reduction) bit-identical to the single call."""
    s = Shape(4, 28, 28, 64, 40, 3, 3, 1, 1, 1, 1, 1, 1)
    s2 = Shape(2, 13, 12, 40, 64, 3, 3, 1, 1, 2, 2, 1, 1)
    cg.force_config(3, splits)  # 64x64, BK 32
    try:
        for layout in (NHWC, NCHW):
            rng = np.random.default_rng(900 + splits + layout)
            X, W, dY, epi = _case(rng, s, layout)
            print(f"deep split-K {splits} layout={layout} {s}", flush=True)
            dW = run_wgrad(layout, s, dY, X)
            assert cg.last_plan()[:2] == (3, splits)
            assert rel_err(dW, oracle.wgrad(layout, s, dY, X))[0] <= TOL
        rng = np.random.default_rng(950 + splits)
        X, W, dY, epi = _case(rng, s, NHWC)
        assert rel_err(run_fwd(NHWC, s, W, X, epi), oracle.fwd(NHWC, s, W, X, **epi))[0] <= TOL
        X2, W2, dY2, _ = _case(rng, s2, NHWC)
        dX0 = rng.standard_normal(s2.x_shape(NHWC)).astype(np.float32)
        got = run_dgrad(NHWC, s2, W2, dY2, dX0)
        assert rel_err(got, oracle.dgrad(NHWC, s2, W2, dY2, dX0=dX0))[0] <= TOL
        # the grouped reduction of a batch sums in the same order as the single launch
        d = cg.make_desc(cg.WGRAD, NHWC, s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr)
        dYd, Xd = dev(dY), dev(X)
        single = torch.empty(s.w_shape(NHWC), device=DEV)
        cg.conv_ex(d, dYd, Xd, single)
        batched = torch.full(s.w_shape(NHWC), float("nan"), device=DEV)
        b = cg.Batch([cg.problem(d, dYd, Xd, batched)], cg.F32, cg.BATCH_GRAPH)
        try:
            b.run()
            torch.cuda.synchronize()
        finally:
            b.close()
        assert torch.equal(single, batched)
    finally:
        cg.force_config(-1, 0)


@pytest.mark.parametrize("layout", [NHWC, NCHW])
@pytest.mark.parametrize("combo", range(8))
def test_epilogue_null_combinations(layout, combo):
    """Any subset of bias / BN / ReLU; NULL means absent (a superset of the
    reference, whose NHWC path dereferences a NULL bias: im2row_nhwc.c:362-371)."""
    s = Shape(2, 9, 9, 8, 12, 3, 3, 1, 1, 1, 1, 1, 1)
    rng = np.random.default_rng(50 + combo)
    X, W, _, epi = _case(rng, s, layout)
    use_bias, use_bn, relu = combo & 1, combo & 2, combo & 4
    e = dict(bias=epi["bias"] if use_bias else None, relu=bool(relu))
    if use_bn:
        e.update(mean=epi["mean"], inv_std=epi["inv_std"], gamma=epi["gamma"], bnbeta=epi["bnbeta"])
    mx, _ = rel_err(run_fwd(layout, s, W, X, e), oracle.fwd(layout, s, W, X, **e))
    assert mx <= TOL


@pytest.mark.parametrize("layout", [NHWC, NCHW])
def test_alpha_beta_extended_api(layout):
    s = Shape(2, 10, 9, 12, 16, 3, 3, 1, 1, 1, 1, 1, 1)
    rng = np.random.default_rng(77)
    X, W, dY, epi = _case(rng, s, layout)
    # forward: Y = epi(alpha*conv + beta*Y0)
    Y0 = rng.standard_normal(s.y_shape(layout)).astype(np.float32)
    Yd = dev(Y0)
    d = cg.make_desc(cg.FWD, layout, s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr)
    cg.conv_ex(d, dev(W), dev(X), Yd, alpha=-0.5, beta=1.5, bias=dev(epi["bias"]), relu=True)
    ref = oracle.fwd(layout, s, W, X, alpha=-0.5, beta=1.5, Y0=Y0, bias=epi["bias"], relu=True)
    assert rel_err(host(Yd), ref)[0] <= TOL
    # weight gradient with beta
    W0 = rng.standard_normal(s.w_shape(layout)).astype(np.float32)
    Wd = dev(W0)
    d = cg.make_desc(cg.WGRAD, layout, s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr)
    cg.conv_ex(d, dev(dY), dev(X), Wd, alpha=2.0, beta=-1.0)
    assert rel_err(host(Wd), oracle.wgrad(layout, s, dY, X, alpha=2.0, beta=-1.0, dW0=W0))[0] <= TOL
    # backward data with beta = 0 (overwrite) and beta = 0.25
    for beta in (0.0, 0.25):
        X0 = rng.standard_normal(s.x_shape(layout)).astype(np.float32)
        Xd = dev(X0)
        d = cg.make_desc(cg.DGRAD, layout, s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr)
        cg.conv_ex(d, dev(W), dev(dY), Xd, alpha=0.75, beta=beta)
        assert rel_err(host(Xd), oracle.dgrad(layout, s, W, dY, alpha=0.75, beta=beta, dX0=X0))[0] <= TOL
    torch.cuda.synchronize()


@pytest.mark.parametrize("op", [cg.FWD, cg.WGRAD, cg.DGRAD])
def test_alpha_zero_scales_output_only(op):
    """quick_gemm.h: alpha == 0 gives C = beta*C without the epilogue."""
    s = Shape(1, 6, 6, 4, 8, 3, 3, 1, 1, 1, 1, 1, 1)
    shape = {cg.FWD: s.y_shape(NHWC), cg.WGRAD: s.w_shape(NHWC), cg.DGRAD: s.x_shape(NHWC)}[op]
    O0 = np.random.default_rng(5).standard_normal(shape).astype(np.float32)
    Od = dev(O0)
    d = cg.make_desc(op, NHWC, s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, 1, 1)
    dummy = dev(np.ones(4096, np.float32))
    cg.conv_ex(d, dummy, dummy, Od, alpha=0.0, beta=-2.0, bias=dummy, relu=True)
    np.testing.assert_array_equal(host(Od), -2.0 * O0)
    cg.conv_ex(d, dummy, dummy, Od, alpha=0.0, beta=0.0)
    assert not host(Od).any()


def test_empty_and_zero_channel_problems_leave_output_untouched():
    Y0 = np.random.default_rng(6).standard_normal((1, 4, 4, 3)).astype(np.float32)
    Yd = dev(Y0)
    dummy = dev(np.ones(64, np.float32))
    # c == 0: the reference's K loop runs zero times and C is not written
    cg.sconvGemmNHWC("N", 1, 4, 4, 0, 3, 3, 3, 1, 1, 1, 1, 1, 1, dummy, dummy, Yd)
    np.testing.assert_array_equal(host(Yd), Y0)
    # b == 0: quick return
    cg.sconvGemmNHWC("N", 0, 4, 4, 2, 3, 3, 3, 1, 1, 1, 1, 1, 1, dummy, dummy, Yd)
    np.testing.assert_array_equal(host(Yd), Y0)


@pytest.mark.parametrize("layout", [NHWC, NCHW])
def test_deterministic_bitwise(layout):
    """No atomics anywhere: two runs are bit-identical, also with split-K."""
    s = Shape(4, 14, 14, 64, 64, 3, 3, 1, 1, 1, 1, 1, 1)
    rng = np.random.default_rng(8)
    X, W, dY, epi = _case(rng, s, layout)
    for cfg, splits in ((-1, 0), (3, 4), (3, 0), (0, 0)):
        cg.force_config(cfg, splits)
        a = run_fwd(layout, s, W, X, epi), run_wgrad(layout, s, dY, X), run_dgrad(
            layout, s, W, dY, np.zeros(s.x_shape(layout), np.float32))
        b = run_fwd(layout, s, W, X, epi), run_wgrad(layout, s, dY, X), run_dgrad(
            layout, s, W, dY, np.zeros(s.x_shape(layout), np.float32))
        for u, v in zip(a, b):
            np.testing.assert_array_equal(u, v)
    cg.force_config(-1, 0)


@pytest.mark.parametrize("layout", [NHWC, NCHW])
def test_reference_acceptance_uniform_data(layout):
    """The reference's own check (tests/test_base.h:12-29): U[0,1] inputs, bias+BN+ReLU
    all on (test_base.c:13-15), element-wise |a-b|/|a| <= 1e-5."""
    s = Shape(2, 16, 16, 32, 48, 3, 3, 1, 1, 1, 1, 1, 1)
    rng = np.random.default_rng(9)
    u = lambda *sh: rng.uniform(0, 1, sh).astype(np.float32)
    X, W = u(*s.x_shape(layout)), u(*s.w_shape(layout))
    epi = dict(bias=u(s.kn), mean=u(s.kn), inv_std=u(s.kn), gamma=u(s.kn), bnbeta=u(s.kn), relu=True)
    Y = run_fwd(layout, s, W, X, epi).astype(np.float64)
    ref = oracle.fwd(layout, s, W, X, **epi)
    nz = ref != 0
    assert np.max(np.abs((ref[nz] - Y[nz]) / ref[nz])) <= 1e-5
    assert np.all(Y[~nz] == 0)


def test_pack_buffers_and_caller_workspace():
    """alloc_pack_buffs workspaces, a caller workspace, and a too-small one."""
    s = Shape(2, 12, 12, 32, 40, 3, 3, 1, 1, 1, 1, 1, 1)
    ac, bc = cg.alloc_pack_buffs()
    try:
        rng = np.random.default_rng(10)
        for layout in (NHWC, NCHW):
            X, W, dY, epi = _case(rng, s, layout)
            cg.force_config(3, 4)  # split-K: partials go to Ac_pack, NCHW weights to Bc_pack
            assert rel_err(run_fwd(layout, s, W, X, epi, packs=(ac, bc)), oracle.fwd(layout, s, W, X, **epi))[0] <= TOL
            dX0 = np.zeros(s.x_shape(layout), np.float32)
            assert rel_err(run_dgrad(layout, s, W, dY, dX0, packs=(ac, bc)), oracle.dgrad(layout, s, W, dY))[0] <= TOL
            cg.force_config(-1, 0)
        d = cg.make_desc(cg.FWD, NCHW, s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, 1, 1)
        need = cg.workspace_size(d)
        ws = torch.empty(need // 4 + 64, device=DEV)
        X, W, _, epi = _case(rng, s, NCHW)
        Yd = torch.empty(s.y_shape(NCHW), device=DEV)
        cg.conv_ex(d, dev(W), dev(X), Yd, workspace=ws, ws_bytes=ws.numel() * 4)
        assert rel_err(host(Yd), oracle.fwd(NCHW, s, W, X))[0] <= TOL
        with pytest.raises(cg.ConvGemmError) as e:
            cg.conv_ex(d, dev(W), dev(X), Yd, workspace=ws, ws_bytes=64)
        assert e.value.status == 4
    finally:
        cg.free_pack_buffs(ac, bc)
        cg.force_config(-1, 0)


def test_dropin_call_sequence():
    """The inferred caller's sequence through bare ctypes declarations (no wrappers):
    alloc_pack_buffs -> sconvGemmNHWC 'N' / 'T' -> sconvGemmNHWC_back -> the NCHW
    three -> convgemm_free_pack_buffs, checked against the oracle."""
    import ctypes
    lib = ctypes.CDLL(cg.LIB_PATH)
    vp, ci = ctypes.c_void_p, ctypes.c_int
    lib.alloc_pack_buffs.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(vp)]
    lib.alloc_pack_buffs.restype = ci
    lib.sconvGemmNHWC.argtypes = [ctypes.c_char] + [ci] * 13 + [vp] * 8 + [ctypes.c_bool, vp, vp]
    lib.sconvGemmNHWC.restype = None
    lib.sconvGemmNCHW.argtypes = lib.sconvGemmNHWC.argtypes
    lib.sconvGemmNCHW.restype = None
    lib.sconvGemmNHWC_back.argtypes = [ci] * 13 + [vp] * 5
    lib.sconvGemmNHWC_back.restype = None
    lib.sconvGemmNCHW_back.argtypes = lib.sconvGemmNHWC_back.argtypes
    lib.sconvGemmNCHW_back.restype = None
    lib.convgemm_free_pack_buffs.argtypes = [vp, vp]
    lib.convgemm_free_pack_buffs.restype = None
    lib.convgemm_get_last_error.restype = ci
    ac, bc = vp(), vp()
    assert lib.alloc_pack_buffs(ctypes.byref(ac), ctypes.byref(bc)) == 0
    s = Shape(2, 10, 9, 16, 24, 3, 3, 1, 1, 2, 1, 1, 1)
    rng = np.random.default_rng(42)
    g = (s.b, s.h, s.w, s.c)
    conv = (s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil, s.hdil)
    back = (s.kn, s.kh, s.kw, s.vstr, s.hstr, s.vpad, s.hpad, s.vdil, s.hdil)  # stride before padding
    for layout in (NHWC, NCHW):
        X, W, dY, epi = _case(rng, s, layout)
        dims = g if layout == NHWC else (s.b, s.c, s.h, s.w)
        fwd = lib.sconvGemmNHWC if layout == NHWC else lib.sconvGemmNCHW
        bwd = lib.sconvGemmNHWC_back if layout == NHWC else lib.sconvGemmNCHW_back
        dW_, dX_, dY_ = dev(W), dev(X), dev(dY)
        e = {k: dev(v) for k, v in epi.items() if k != "relu"}
        Y = torch.empty(s.y_shape(layout), device=DEV)
        fwd(b"N", *dims, *conv, dW_.data_ptr(), dX_.data_ptr(), Y.data_ptr(), e["bias"].data_ptr(),
            e["mean"].data_ptr(), e["inv_std"].data_ptr(), e["gamma"].data_ptr(), e["bnbeta"].data_ptr(),
            bool(epi["relu"]), ac, bc)
        assert lib.convgemm_get_last_error() == 0
        assert rel_err(host(Y), oracle.fwd(layout, s, W, X, **epi))[0] <= TOL
        gW = torch.empty(s.w_shape(layout), device=DEV)
        fwd(b"T", *dims, *conv, dY_.data_ptr(), dX_.data_ptr(), gW.data_ptr(), None, None, None, None, None,
            False, ac, bc)
        assert rel_err(host(gW), oracle.wgrad(layout, s, dY, X))[0] <= TOL
        gX = torch.zeros(s.x_shape(layout), device=DEV)
        bwd(*dims, *back, dW_.data_ptr(), dY_.data_ptr(), gX.data_ptr(), ac, bc)
        assert lib.convgemm_get_last_error() == 0
        assert rel_err(host(gX), oracle.dgrad(layout, s, W, dY))[0] <= TOL
    lib.convgemm_free_pack_buffs(ac, bc)


def _fixture_shapes(max_macs=6e7, limit=40):
    """Unique CNN layer shapes from the reference's shape list (tests/test.txt,
    columns h w c b kh kw kn stride padding), at batch 1."""
    rows = []
    seen = set()
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "cnn_layer_shapes.txt")
    for line in open(path):
        if line.startswith("#") or not line.strip():
            continue
        h, w, c, b, kh, kw, kn, st, pad = map(int, line.split())
        key = (h, w, c, kh, kw, kn, st, pad)
        if key in seen:
            continue
        seen.add(key)
        s = Shape(1, h, w, c, kn, kh, kw, pad, pad, st, st, 1, 1)
        if s.ho <= 0 or s.wo <= 0 or s.P * s.K * s.kn > max_macs:
            continue
        rows.append(s)
    rows.sort(key=lambda s: s.P * s.K * s.kn)
    step = max(1, len(rows) // limit)
    return rows[::step][:limit]


@pytest.mark.parametrize("si", range(len(_fixture_shapes())))
def test_cnn_layer_shape_fixture(si):
    s = _fixture_shapes()[si]
    for layout in (NHWC, NCHW):
        rng = np.random.default_rng(4000 + si)
        X, W, dY, epi = _case(rng, s, layout)
        assert rel_err(run_fwd(layout, s, W, X, epi), oracle.fwd(layout, s, W, X, **epi))[0] <= TOL
        assert rel_err(run_wgrad(layout, s, dY, X), oracle.wgrad(layout, s, dY, X))[0] <= TOL
        dX0 = np.zeros(s.x_shape(layout), np.float32)
        assert rel_err(run_dgrad(layout, s, W, dY, dX0), oracle.dgrad(layout, s, W, dY))[0] <= TOL


SK_SHAPES = [
    Shape(4, 14, 14, 64, 64, 3, 3, 1, 1, 1, 1, 1, 1),    # several tiles per workgroup range
    Shape(2, 9, 7, 36, 20, 3, 3, 1, 1, 2, 1, 1, 1),      # fewer units than workgroups
    Shape(8, 28, 28, 32, 96, 1, 1, 0, 0, 1, 1, 1, 1),    # 1x1, ragged N tiles
    Shape(3, 16, 16, 40, 24, 5, 5, 2, 2, 1, 1, 1, 1),    # long K per tile
]


@pytest.mark.parametrize("cfg", [0, 3, 5, 6, 11])
@pytest.mark.parametrize("si", range(len(SK_SHAPES)))
def test_stream_k(cfg, si):
    """Stream-K (splits 0): forward with the full epilogue and weight gradient,
    both layouts, against the oracle; the plan reports splits 0 and two kernels."""
    s = SK_SHAPES[si]
    cg.force_config(cfg, 0)
    try:
        for layout in (NHWC, NCHW):
            rng = np.random.default_rng(900 + 10 * si + cfg)
            X, W, dY, epi = _case(rng, s, layout)
            y = run_fwd(layout, s, W, X, epi)
            plan = cg.last_plan()
            assert plan[1] == 0 and plan[2] >= 2, plan
            assert rel_err(y, oracle.fwd(layout, s, W, X, **epi))[0] <= TOL
            dw = run_wgrad(layout, s, dY, X)
            assert cg.last_plan()[1] == 0
            assert rel_err(dw, oracle.wgrad(layout, s, dY, X))[0] <= TOL
    finally:
        cg.force_config(-1, 0)


@pytest.mark.parametrize("dtype", ["f32", "bf16", "fp8"])
def test_prepacked_weights_bit_identical(dtype):
    """pack_weights once + conv_packed gives the bits of conv_ex (same plan), for
    every operation that repacks (forward, NCHW backward data) and the ones that
    read weights as given (packed size 0); forced through every pipelined forward
    configuration, conv_ex reads the caller's weights in place and conv_packed
    the packed ones."""
    elem = {"f32": cg.F32, "bf16": cg.BF16, "fp8": cg.FP8_E4M3}[dtype]
    s = Shape(2, 12, 11, 32, 40, 3, 3, 1, 1, 2, 1, 1, 1)
    g = torch.Generator().manual_seed(77)
    tdt = {"f32": torch.float32, "bf16": torch.bfloat16, "fp8": torch.float8_e4m3fn}[dtype]

    def operand(shape):
        t = torch.randn(shape, generator=g).to(tdt)
        if dtype == "bf16":
            t = t.view(torch.int16)
        elif dtype == "fp8":
            t = t.view(torch.uint8)
        return t.to(DEV)

    scales = (0.5, 2.0) if dtype == "fp8" else None
    ops = [(cg.FWD, NHWC), (cg.FWD, NCHW)] + ([] if dtype == "fp8" else [(cg.DGRAD, NHWC), (cg.DGRAD, NCHW)])
    for op, layout in ops:
        d = cg.make_desc(op, layout, s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr)
        W = operand(s.w_shape(layout))
        Xin = operand(s.x_shape(layout) if op == cg.FWD else s.y_shape(layout))
        out_shape = s.y_shape(layout) if op == cg.FWD else s.x_shape(layout)
        y0 = torch.zeros(out_shape, device=DEV)
        y1 = torch.zeros(out_shape, device=DEV)
        beta = 1.0 if op == cg.DGRAD else 0.0
        cg.conv_ex(d, W, Xin, y0, beta=beta, dtype=elem, scales=scales)
        torch.cuda.synchronize()
        plan = cg.last_plan()
        n = cg.packed_weights_size(d, elem)
        packed = torch.empty(max(n, 1), dtype=torch.uint8, device=DEV) if n else W
        if n:
            cg.pack_weights(d, W, packed, dtype=elem)
        cg.conv_packed(d, packed, Xin, y1, beta=beta, dtype=elem, scales=scales)
        torch.cuda.synchronize()
        kplan = cg.last_plan()
        # a pipelined or halo-tile forward NHWC plan whose KN fills 16-byte chunks reads the caller's
        # weights in place: conv_ex runs no repack either, and the kernel reads the other weight order
        # with the same MFMA inputs
        t = cg.tile_configs(elem)[plan[0]]
        es = {"f32": 4, "bf16": 2, "fp8": 1}[dtype]
        raw = op == cg.FWD and layout == NHWC and t[5] > 0 and (s.kn * es) % 16 == 0
        assert kplan[:2] == plan[:2] and kplan[2] == plan[2] - (1 if n and not raw else 0), (plan, kplan, raw)
        assert torch.equal(y0, y1), (op, layout)


INPLACE_SHAPES = [
    Shape(2, 12, 11, 64, 80, 3, 3, 1, 1, 2, 1, 1, 1),   # N tail, stride
    Shape(1, 9, 9, 160, 48, 3, 3, 1, 1, 1, 1, 1, 1),    # ragged last channel block
    Shape(3, 7, 6, 32, 16, 1, 1, 0, 0, 1, 1, 1, 1),     # M tail, one K step
]


@pytest.mark.parametrize("dtype", ["f32", "bf16", "fp8"])
def test_inplace_weights_every_pipelined_config(dtype):
    """Every pipelined and halo-tile forward NHWC configuration, split-K 1 and 2: conv_ex
    reads the caller's [c][kh][kw][kn] weights in place (transposing LDS reads, no repack
    kernel) and gives the bits of conv_packed, which reads them repacked to [kn][kh][kw][c].
    A halo configuration whose halo does not fit the shape is replaced by the planner."""
    elem = {"f32": cg.F32, "bf16": cg.BF16, "fp8": cg.FP8_E4M3}[dtype]
    tdt = {"f32": torch.float32, "bf16": torch.bfloat16, "fp8": torch.float8_e4m3fn}[dtype]
    view = {"f32": torch.float32, "bf16": torch.int16, "fp8": torch.uint8}[dtype]
    scales = (0.5, 2.0) if dtype == "fp8" else None
    confs = cg.tile_configs(elem)
    lp = [i for i, t in enumerate(confs) if t[5] > 0 and t[6] == 0 and cg.tile_supports(elem, i, cg.FWD, NHWC)]
    assert lp and any(confs[i][7] for i in lp)
    g = torch.Generator().manual_seed(91)
    halo_served = 0
    for si, s in enumerate(INPLACE_SHAPES):
        d = cg.make_desc(cg.FWD, NHWC, s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr)
        W = (torch.randn(s.w_shape(NHWC), generator=g) * 0.5).to(tdt).view(view).to(DEV)
        X = torch.randn(s.x_shape(NHWC), generator=g).to(tdt).view(view).to(DEV)
        packed = torch.empty(cg.packed_weights_size(d, elem), dtype=torch.uint8, device=DEV)
        cg.pack_weights(d, W, packed, dtype=elem)
        for cfg in lp:
            for splits in (1, 2):
                print(f"in-place {dtype} shape {si} cfg {cfg} splits {splits}", flush=True)
                y0 = torch.full(s.y_shape(NHWC), float("nan"), device=DEV)
                y1 = torch.full(s.y_shape(NHWC), float("nan"), device=DEV)
                cg.force_config(cfg, splits)
                try:
                    cg.conv_ex(d, W, X, y0, dtype=elem, scales=scales)
                    torch.cuda.synchronize()
                    plan = cg.last_plan()
                    cg.conv_packed(d, packed, X, y1, dtype=elem, scales=scales)
                    torch.cuda.synchronize()
                    kplan = cg.last_plan()
                finally:
                    cg.force_config(-1, 0)
                assert (plan[0] == cfg or confs[cfg][7]) and kplan[:3] == plan[:3], (plan, kplan)
                assert torch.equal(y0, y1), (si, cfg, splits)
                halo_served += plan[0] == cfg and confs[cfg][7] > 0
    assert halo_served, "no halo configuration served a shape"
