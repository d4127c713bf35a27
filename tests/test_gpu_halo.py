"""The halo-tile forward kernel (NHWC, FP32 / BF16 / FP8 operands).

Its M tiles are runs of output pixels of one image, and the input rows those
pixels read (the halo) are loaded into LDS once per channel chunk; the taps
read their operands from it at shifted offsets.  Checks:
* every halo configuration, forced on shapes with padding, stride, dilation,
  5x5 and 7x7 kernels, images smaller than a tile, tiles of several output
  rows and of pieces of a row, channel counts that end inside a 128-byte chunk, with split-K 1-3
  and the fused epilogue, against the FP64 oracle on the dequantized operands;
* the plan falls back to another kernel exactly when the largest halo of the
  problem does not fit the configuration's slots (counted here by brute force
  over the input rows each tile touches);
* alpha/beta accumulation, grouped launches (bit-identical to single calls),
  and the ResNet-50 stride-1 3x3 layers at batch 32 (sampled FP64 oracle).
"""
import numpy as np
import pytest
import torch

import oracle
from oracle import NHWC, Shape, sampled

pytestmark = pytest.mark.gpu

import convgemm_amd as cg  # noqa: E402
from tests.gpu_util import DEV, rel_err  # noqa: E402
from tests.test_gpu_lowp import FP8_TOL, q_bf16, q_fp8  # noqa: E402

ELEM = {"f32": cg.F32, "bf16": cg.BF16, "fp8": cg.FP8_E4M3}
TOL = {"f32": 1e-5, "bf16": 2e-5, "fp8": FP8_TOL}
CMULT = {"f32": 4, "bf16": 8, "fp8": 16}  # channels per 16 bytes

SHAPES = [
    Shape(2, 12, 11, 32, 48, 3, 3, 1, 1, 1, 1, 1, 1),
    Shape(2, 9, 9, 48, 40, 3, 3, 1, 0, 2, 2, 1, 1),     # stride 2, asymmetric padding
    Shape(2, 10, 10, 80, 24, 3, 3, 2, 2, 1, 1, 2, 2),   # dilation 2; channels end inside a chunk
    Shape(1, 17, 23, 64, 72, 5, 5, 2, 2, 1, 1, 1, 1),   # 5x5, kn not a multiple of the tile
    Shape(3, 7, 7, 128, 96, 3, 3, 1, 1, 1, 1, 1, 1),    # images smaller than one M tile
    Shape(2, 11, 11, 16, 40, 7, 7, 3, 3, 2, 2, 1, 1),   # 7x7 stride 2: fits the larger halos only
    Shape(2, 6, 30, 24, 32, 3, 3, 0, 1, 1, 1, 1, 1),    # wide rows, no vertical padding
    Shape(4, 28, 28, 64, 64, 3, 3, 1, 1, 1, 1, 1, 1),   # several tiles per image, split rows
    Shape(2, 20, 20, 256, 32, 3, 3, 1, 1, 1, 1, 1, 1),  # many channel chunks
    Shape(1, 5, 150, 16, 24, 3, 3, 1, 1, 1, 1, 1, 1),   # rows wider than a tile: pieces of a row
    Shape(1, 6, 301, 32, 16, 1, 3, 0, 1, 2, 2, 1, 1),   # 1x3, stride 2, pieces of a row
]


def halo_cfgs(elem):
    """Halo configurations in raster order 0: (index, BM, halo slots)."""
    ns = cg.num_tile_shapes(elem)
    return [(i, t[0], t[7]) for i, t in enumerate(cg.tile_configs(elem)[:ns]) if t[7] > 0]


def halo_tiles(s: Shape, bm):
    """Output pixels (indices within an image) of each halo tile: whole output
    rows when a row fits the bm tile rows, else bm-column pieces of a row."""
    if s.wo <= bm:
        r = bm // s.wo
        return [np.arange(r0 * s.wo, min(r0 + r, s.ho) * s.wo) for r0 in range(0, s.ho, r)]
    return [np.arange(row * s.wo + c0, row * s.wo + min(c0 + bm, s.wo))
            for row in range(s.ho) for c0 in range(0, s.wo, bm)]


def halo_fits(s: Shape, bm, slots):
    """Largest halo over the tiles of an image, counted by listing the input
    rows and columns (padding included) the tile's pixels read."""
    if s.kh * s.kw == 1:
        return False
    worst = 0
    for qs in halo_tiles(s, bm):
        assert 0 < len(qs) <= bm
        rows = (qs // s.wo)[:, None] * s.vstr + np.arange(s.kh)[None, :] * s.vdil
        cols = (qs % s.wo)[:, None] * s.hstr + np.arange(s.kw)[None, :] * s.hdil
        worst = max(worst, (rows.max() - rows.min() + 1) * (cols.max() - cols.min() + 1))
    return worst <= slots


def _operands(dtype, X, W):
    if dtype == "f32":
        return torch.from_numpy(X).to(DEV), X, torch.from_numpy(W).to(DEV), W, None
    if dtype == "bf16":
        (Xd, Xq), (Wd, Wq) = q_bf16(X), q_bf16(W)
        return Xd, Xq, Wd, Wq, None
    Xd, Xq, sx = q_fp8(X)
    Wd, Wq, sw = q_fp8(W)
    return Xd, Xq, Wd, Wq, (sw, sx)


def _geo(s):
    return (s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil, s.hdil)


@pytest.mark.parametrize("dtype", ["f32", "bf16", "fp8"])
@pytest.mark.parametrize("si", range(len(SHAPES)))
def test_halo_configurations(dtype, si):
    elem = ELEM[dtype]
    cfgs = halo_cfgs(elem)
    assert cfgs, "no halo configurations"
    s = SHAPES[si]
    if s.c % CMULT[dtype]:
        pytest.skip("16-byte channel chunks need C * element size % 16 == 0")
    rng = np.random.default_rng(90 + si)
    X = rng.standard_normal(s.x_shape(NHWC)).astype(np.float32)
    W = (rng.standard_normal(s.w_shape(NHWC)) * 0.1).astype(np.float32)
    Xd, Xq, Wd, Wq, scales = _operands(dtype, X, W)
    epi = dict(bias=(0.1 * rng.standard_normal(s.kn)).astype(np.float32), relu=True)
    ref = oracle.fwd(NHWC, s, Wq, Xq, **epi)
    d = cg.make_desc(cg.FWD, NHWC, *_geo(s))
    bias = torch.from_numpy(epi["bias"]).to(DEV)
    served = 0
    for cfg, bm, slots in cfgs:
        fits = halo_fits(s, bm, slots)
        for splits in (1, 2, 3):
            cg.force_config(cfg, splits)
            try:
                Y = torch.full(s.y_shape(NHWC), float("nan"), device=DEV)
                print(f"halo {dtype} shape {si} cfg {cfg} splits {splits} fits {fits}", flush=True)
                cg.conv_ex(d, Wd, Xd, Y, dtype=elem, scales=scales, bias=bias, relu=True)
                torch.cuda.synchronize()
                used = cg.last_plan()[0]
            finally:
                cg.force_config(-1, 0)
            assert (used == cfg) == fits, (cfg, used, fits)
            served += used == cfg
            assert rel_err(Y.cpu().numpy(), ref)[0] <= TOL[dtype], (cfg, splits)
    assert served, "no halo configuration served the shape"


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
def test_halo_alpha_beta_full_epilogue(dtype):
    """Y = alpha * conv + beta * Y0 followed by bias, batch norm and ReLU."""
    elem = ELEM[dtype]
    s = Shape(2, 14, 14, 64, 80, 3, 3, 1, 1, 1, 1, 1, 1)
    rng = np.random.default_rng(7)
    X = rng.standard_normal(s.x_shape(NHWC)).astype(np.float32)
    W = (rng.standard_normal(s.w_shape(NHWC)) * 0.1).astype(np.float32)
    Y0 = rng.standard_normal(s.y_shape(NHWC)).astype(np.float32)
    Xd, Xq, Wd, Wq, scales = _operands(dtype, X, W)
    kn = s.kn
    epi = dict(bias=0.1 * rng.standard_normal(kn), mean=0.1 * rng.standard_normal(kn),
               inv_std=1.0 / np.sqrt(0.5 + rng.random(kn)), gamma=0.5 + rng.random(kn),
               bnbeta=0.1 * rng.standard_normal(kn))
    epi = {k: v.astype(np.float32) for k, v in epi.items()}
    ref = oracle.fwd(NHWC, s, Wq, Xq, alpha=0.75, beta=0.5, Y0=Y0, relu=True, **epi)
    d = cg.make_desc(cg.FWD, NHWC, *_geo(s))
    t = {k: torch.from_numpy(v).to(DEV) for k, v in epi.items()}
    for cfg, bm, slots in halo_cfgs(elem):
        assert halo_fits(s, bm, slots)
        for splits in (1, 2):
            cg.force_config(cfg, splits)
            try:
                Y = torch.from_numpy(Y0).to(DEV)
                cg.conv_ex(d, Wd, Xd, Y, alpha=0.75, beta=0.5, dtype=elem, scales=scales, relu=True, **t)
                torch.cuda.synchronize()
                assert cg.last_plan()[0] == cfg
            finally:
                cg.force_config(-1, 0)
            assert rel_err(Y.cpu().numpy(), ref)[0] <= TOL[dtype], (cfg, splits)


@pytest.mark.parametrize("dtype", ["f32", "bf16", "fp8"])
@pytest.mark.parametrize("flags", [cg.BATCH_GRAPH, cg.BATCH_GRAPH | cg.BATCH_PERSISTENT])
def test_halo_groups(dtype, flags):
    """Grouped (and persistent) launches of each halo configuration on uneven
    3x3 problems: bit-identical to one call per problem."""
    from tests.test_gpu_batch import BN_RELU, Case, D, check
    elem = ELEM[dtype]
    for cfg, bm, slots in halo_cfgs(elem):
        for splits in (1, 2):
            rng = np.random.default_rng(500 + cfg * 7 + splits)
            cases = []
            while len(cases) < 5:
                i = len(cases)
                b, h = int(rng.choice([1, 2, 3])), int(rng.choice([6, 9, 12]))
                st = int(rng.choice([1, 2]))
                c = int(rng.choice(range(16, 161, 16)))
                kn = int(rng.choice(range(16, 161, 16)))
                if not halo_fits(Shape(b, h, h, c, kn, 3, 3, 1, 1, st, st, 1, 1), bm, slots):
                    continue
                cases.append(Case(D(cg.FWD, NHWC, b, h, h, c, kn, 3, s=st), elem,
                                  600 + i, alpha=0.75 if i % 4 == 1 else 1.0, beta=1.0 if i % 4 == 3 else 0.0,
                                  epi=BN_RELU if i % 2 == 0 else None))
            print(f"halo group {dtype} cfg {cfg} splits {splits} flags {flags}", flush=True)
            cg.force_config(cfg, splits)
            try:
                refs = []
                for c in cases:
                    refs.append(c.reference())
                    assert cg.last_plan()[0] == cfg
                b = cg.Batch([c.problem() for c in cases], elem, flags)
            finally:
                cg.force_config(-1, 0)
            for rep in range(2):
                for c in cases:
                    c.reset()
                b.run()
                torch.cuda.synchronize()
                check(cases, refs, f"cfg={cfg} splits={splits} replay={rep}")
            b.close()


LAYERS = {
    "L2_3x3_64to64_56": Shape(32, 56, 56, 64, 64, 3, 3, 1, 1, 1, 1, 1, 1),
    "L5_3x3_512to512_7": Shape(32, 7, 7, 512, 512, 3, 3, 1, 1, 1, 1, 1, 1),
}


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
@pytest.mark.parametrize("name", list(LAYERS))
def test_halo_fullsize(name, dtype):
    """Each halo configuration whose halo fits, forced on the full-size layer
    with bias + ReLU (split-K 1 and 2): 256 sampled outputs against the FP64
    oracle on the dequantized operands."""
    s = LAYERS[name]
    elem = ELEM[dtype]
    g = torch.Generator(device="cuda").manual_seed(sum(map(ord, name)) + 5)
    X = torch.randn(s.x_shape(NHWC), device="cuda", generator=g)
    W = torch.randn(s.w_shape(NHWC), device="cuda", generator=g) * (2.0 / s.K) ** 0.5
    bias = 0.1 * torch.randn(s.kn, device="cuda", generator=g)
    if dtype == "bf16":
        Xd = torch.empty(X.shape, dtype=torch.int16, device=DEV)
        Wd = torch.empty(W.shape, dtype=torch.int16, device=DEV)
        cg.quantize(X, Xd, cg.BF16)
        cg.quantize(W, Wd, cg.BF16)
        Xh = Xd.cpu().view(torch.bfloat16).double().numpy()
        Wh = Wd.cpu().view(torch.bfloat16).double().numpy()
    else:
        Xd, Wd, Xh, Wh = X, W, X.cpu().numpy(), W.cpu().numpy()
    rng = np.random.default_rng(12)
    ys = np.stack([rng.integers(0, n, 256) for n in (s.b, s.ho, s.wo, s.kn)], 1)
    ref = sampled.fwd_sampled(NHWC, s, Wh, Xh, ys, bias=bias.cpu().numpy(), relu=True)
    d = cg.make_desc(cg.FWD, NHWC, *_geo(s))
    n = 0
    for cfg, bm, slots in halo_cfgs(elem):
        if not halo_fits(s, bm, slots):
            continue
        for splits in (1, 2):
            print(f"halo full-size {name} {dtype} cfg {cfg} splits {splits}", flush=True)
            cg.force_config(cfg, splits)
            try:
                Y = torch.full(s.y_shape(NHWC), float("nan"), device=DEV)
                cg.conv_ex(d, Wd, Xd, Y, dtype=elem, bias=bias, relu=True)
                torch.cuda.synchronize()
                assert cg.last_plan()[0] == cfg
            finally:
                cg.force_config(-1, 0)
            got = sampled.nhwc_index(NHWC, Y.cpu().numpy(), ys)
            assert np.max(np.abs(got - ref)) <= TOL[dtype] * max(np.max(np.abs(ref)), 1.0), (cfg, splits)
            n += 1
    assert n
