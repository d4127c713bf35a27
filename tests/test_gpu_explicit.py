"""Explicit path on the GPU: transforms, plain GEMM, layout converters.

* im2row_nhwc / im2col_nchw are copies: bit-exact against the FP64 oracle
  transforms (oracle/conv_oracle.c), and padded taps are left untouched (a
  sentinel survives), as in the reference's active branch.
* row2im_nhwc / col2im_nchw accumulate into the image: the FP64 oracle sum
  within 1e-6 relative (FP32 sums of at most kh*kw terms).
* transpose_nchw: bit-exact against the oracle's swap01, with beta and with
  start_row / start_col offsets.
* convgemm_gemm: every order x transpose combination with padded leading
  dimensions against an FP64 product, max-rel <= 1e-5 (FP32), BF16 and FP8
  against the product of the dequantized operands, and the quick-return
  rules of the reference.
* explicit path (im2row/im2col + GEMM [+ transpose_nchw]) against the
  implicit convolution: the same convolution within 1e-5.
"""
import itertools

import numpy as np
import pytest
import torch

import oracle
from oracle import NCHW, NHWC, Shape

pytestmark = pytest.mark.gpu

import convgemm_amd as cg  # noqa: E402
from tests.gpu_util import DEV, dev, host, rel_err  # noqa: E402

SHAPES = [
    Shape(2, 7, 9, 5, 6, 3, 3, 1, 1, 1, 1, 1, 1),
    Shape(1, 11, 10, 4, 3, 3, 2, 2, 0, 2, 3, 1, 1),
    Shape(2, 9, 9, 3, 5, 3, 3, 2, 2, 1, 1, 2, 2),
    Shape(3, 6, 5, 8, 4, 1, 1, 0, 0, 1, 1, 1, 1),
    Shape(1, 13, 12, 6, 7, 5, 3, 2, 1, 2, 2, 1, 2),
]


def targs(s):
    return (s.ho, s.wo, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil, s.hdil)


@pytest.mark.parametrize("si", range(len(SHAPES)))
def test_im2row_im2col_exact(si):
    s = SHAPES[si]
    rng = np.random.default_rng(si)
    for layout in (NHWC, NCHW):
        X = rng.standard_normal(s.x_shape(layout)).astype(np.float32)
        if layout == NHWC:
            ref = oracle.im2row_nhwc(s, X)  # P x K, zeros at padded taps
            mask = oracle.im2row_nhwc(s, np.ones_like(X)) != 0
            buf = torch.full((s.P, s.K), 7.0, device=DEV)
            cg.im2row_nhwc(buf, s.K, dev(X), s.b, s.h, s.w, s.c, *targs(s))
        else:
            ref = oracle.im2col_nchw(s, X)  # K x P
            mask = oracle.im2col_nchw(s, np.ones_like(X)) != 0
            buf = torch.full((s.K, s.P), 7.0, device=DEV)
            cg.im2col_nchw(buf, s.P, dev(X), s.b, s.c, s.h, s.w, *targs(s))
        got = host(buf)
        assert np.array_equal(got[mask], ref[mask].astype(np.float32))
        assert np.all(got[~mask] == 7.0)  # padded taps untouched


@pytest.mark.parametrize("dtype", [cg.F32, cg.BF16])
@pytest.mark.parametrize("si", range(len(SHAPES)))
def test_im2row_ex_exact(si, dtype):
    """convgemm_im2row_ex: the same copy in FP32 or BF16 elements, bit-exact against the
    oracle's transform of the same values, padded taps untouched, a padded leading dimension."""
    s = SHAPES[si]
    rng = np.random.default_rng(50 + si)
    for layout in (NHWC, NCHW):
        X = rng.standard_normal(s.x_shape(layout)).astype(np.float32)
        if dtype == cg.BF16:  # values exactly representable in BF16
            X = torch.from_numpy(X).to(torch.bfloat16).float().numpy()
        tdt = torch.float32 if dtype == cg.F32 else torch.bfloat16
        if layout == NHWC:
            ref, mask = oracle.im2row_nhwc(s, X), oracle.im2row_nhwc(s, np.ones_like(X)) != 0
            rows, cols = s.P, s.K
        else:
            ref, mask = oracle.im2col_nchw(s, X), oracle.im2col_nchw(s, np.ones_like(X)) != 0
            rows, cols = s.K, s.P
        ld = cols + 3
        buf = torch.full((rows, ld), 7.0, device=DEV, dtype=tdt)
        x = dev(X).to(tdt)
        d = cg.make_desc(cg.FWD, layout, s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr,
                         s.vdil, s.hdil)
        cg.im2row_ex(d, x.view(torch.int16) if dtype == cg.BF16 else x,
                     buf.view(torch.int16) if dtype == cg.BF16 else buf, ld, dtype=dtype)
        got = host(buf.float())
        assert np.array_equal(got[:, :cols][mask], ref[mask].astype(np.float32)), (si, layout)
        assert np.all(got[:, :cols][~mask] == 7.0) and np.all(got[:, cols:] == 7.0)


@pytest.mark.parametrize("si", range(len(SHAPES)))
def test_row2im_col2im(si):
    s = SHAPES[si]
    rng = np.random.default_rng(100 + si)
    for layout in (NHWC, NCHW):
        X0 = rng.standard_normal(s.x_shape(layout)).astype(np.float32)
        if layout == NHWC:
            R = rng.standard_normal((s.P, s.K)).astype(np.float32)
            ref = oracle.row2im_nhwc(s, R, X0)
            out = dev(X0)
            cg.row2im_nhwc(s.P, s.K, dev(R), s.K, out, s.b, s.h, s.w, s.c, *targs(s))
        else:
            R = rng.standard_normal((s.K, s.P)).astype(np.float32)
            ref = oracle.col2im_nchw(s, R, X0)
            out = dev(X0)
            cg.col2im_nchw(s.K, s.P, dev(R), s.P, out, s.b, s.c, s.h, s.w, *targs(s))
        mx, _ = rel_err(host(out), ref)
        assert mx <= 1e-6, (layout, mx)


@pytest.mark.parametrize("beta", [0.0, 0.5])
def test_transpose_nchw(beta):
    b, kn, ho, wo = 3, 5, 4, 6
    rng = np.random.default_rng(7)
    blk = rng.standard_normal((kn, b, ho, wo)).astype(np.float32)  # GEMM output [kn][b*ho*wo]
    out0 = rng.standard_normal((b, kn, ho, wo)).astype(np.float32)
    ref = oracle.swap01(blk, beta=beta, out0=out0)
    out = dev(out0)
    P = b * ho * wo
    cg.transpose_nchw(P, kn, dev(blk), P, beta, out, kn, ho, wo, 0, 0)
    assert np.array_equal(host(out), ref.astype(np.float32)) if beta == 0.0 else \
        rel_err(host(out), ref)[0] <= 1e-7
    # a sub-block: pixels [r0, r0+nr) of channels [c0, c0+nc), source ld = nr + 3
    r0, nr, c0, nc = ho * wo, 2 * ho * wo, 1, 3  # batches 1..2, channels 1..3
    src = rng.standard_normal((nc, nr + 3)).astype(np.float32)
    out = dev(out0)
    cg.transpose_nchw(nr, nc, dev(src), nr + 3, 0.0, out, kn, ho, wo, r0, c0)
    exp = out0.copy()
    exp[1:3, c0:c0 + nc] = oracle.swap01(src[:, :nr].reshape(nc, 2, ho, wo)).astype(np.float32)
    assert np.array_equal(host(out), exp)


def _stored(rng, order, trans, rows, cols, pad):
    """A matrix op(S) of rows x cols, stored with `order`/`trans` and leading-dim padding.
    Returns (device buffer, ld, the float64 op(S))."""
    opS = rng.standard_normal((rows, cols))
    S = opS if trans == "N" else opS.T  # the stored matrix
    r, c = S.shape
    if order == "R":
        ld = c + pad
        buf = np.zeros((max(r, 1), ld), np.float32)
        buf[:r, :c] = S
    else:
        ld = r + pad
        buf = np.zeros((max(c, 1), ld), np.float32)
        buf[:c, :r] = S.T
    return dev(buf.ravel()), ld, opS.astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("oa,ob,oc,ta,tb", list(itertools.product("CR", "CR", "CR", "NT", "NT")))
def test_gemm_all_layouts(oa, ob, oc, ta, tb):
    m, n, k = 37, 29, 53
    rng = np.random.default_rng(sum(ord(ch) << (5 * i) for i, ch in enumerate(oa + ob + oc + ta + tb)))
    A, lda, a = _stored(rng, oa, ta, m, k, 3)
    B, ldb, b = _stored(rng, ob, tb, k, n, 0 if oc == "R" else 5)
    C0 = rng.standard_normal((m, n))
    ldc = (n if oc == "R" else m) + (2 if ta == "T" else 0)
    if oc == "R":
        cbuf = np.zeros((m, ldc), np.float32); cbuf[:, :n] = C0
    else:
        cbuf = np.zeros((n, ldc), np.float32); cbuf[:, :m] = C0.T
    C = dev(cbuf.ravel())
    alpha, beta = 1.5, -0.5
    cg.gemm(oa, ob, oc, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
    got = host(C).reshape(cbuf.shape)
    got = got[:, :n] if oc == "R" else got[:, :m].T
    ref = alpha * (a @ b) + beta * C0.astype(np.float32)
    mx, _ = rel_err(got, ref)
    assert mx <= 1e-5, mx
    pad = host(C).reshape(cbuf.shape)[:, (n if oc == "R" else m):]
    assert np.all(pad == 0)  # beyond the matrix: untouched


def test_gemm_quick_returns_and_errors():
    rng = np.random.default_rng(3)
    C0 = rng.standard_normal((8, 6)).astype(np.float32)
    A = dev(rng.standard_normal((8, 4)).astype(np.float32))
    B = dev(rng.standard_normal((4, 6)).astype(np.float32))
    C = dev(C0)
    cg.gemm("R", "R", "R", "N", "N", 8, 6, 4, 0.0, A, 4, B, 6, 2.0, C, 6)  # alpha == 0: C = beta*C
    assert np.array_equal(host(C), 2.0 * C0)
    C = dev(C0)
    cg.gemm("R", "R", "R", "N", "N", 8, 6, 0, 1.0, A, 4, B, 6, 3.0, C, 6)  # k == 0: C untouched
    assert np.array_equal(host(C), C0)
    cg.gemm("R", "R", "R", "N", "N", 0, 6, 4, 1.0, A, 4, B, 6, 0.0, C, 6)  # m == 0: nothing
    assert np.array_equal(host(C), C0)
    with pytest.raises(cg.ConvGemmError):
        cg.gemm("X", "R", "R", "N", "N", 8, 6, 4, 1.0, A, 4, B, 6, 0.0, C, 6)
    with pytest.raises(cg.ConvGemmError):
        cg.gemm("R", "R", "R", "N", "N", 8, 6, 4, 1.0, A, 3, B, 6, 0.0, C, 6)  # ldA < k


@pytest.mark.parametrize("dtype", [cg.BF16, cg.FP8_E4M3])
def test_gemm_low_precision(dtype):
    m, n, k = 64, 48, 96
    a = torch.randn(m, k, generator=torch.Generator().manual_seed(1))
    b = torch.randn(k, n, generator=torch.Generator().manual_seed(2))
    if dtype == cg.BF16:
        qa, qb = a.to(torch.bfloat16), b.to(torch.bfloat16)
        A, B = qa.view(torch.int16).to(DEV), qb.view(torch.int16).to(DEV)
        scales, sa, sb = None, 1.0, 1.0
    else:
        qa, qb = a.to(torch.float8_e4m3fn), b.to(torch.float8_e4m3fn)
        A, B = qa.view(torch.uint8).to(DEV), qb.view(torch.uint8).to(DEV)
        sa, sb = 0.5, 2.0
        scales = (sa, sb)
    C = torch.empty(m * n, device=DEV)
    cg.gemm("R", "R", "R", "N", "N", m, n, k, 1.0, A, k, B, n, 0.0, C, n, dtype=dtype, scales=scales)
    ref = (qa.double().numpy() * sa) @ (qb.double().numpy() * sb)
    mx, _ = rel_err(host(C).reshape(m, n), ref)
    assert mx <= (2e-5 if dtype == cg.BF16 else 1e-4), mx


@pytest.mark.parametrize("si", [0, 2, 4])
def test_explicit_equals_implicit(si):
    s = SHAPES[si]
    rng = np.random.default_rng(200 + si)
    # NHWC: Y[P][kn] = rows[P][K] W[K][kn]
    X = rng.standard_normal(s.x_shape(NHWC)).astype(np.float32)
    W = rng.standard_normal(s.w_shape(NHWC)).astype(np.float32)
    rows = torch.zeros(s.P * s.K, device=DEV)
    cg.im2row_nhwc(rows, s.K, dev(X), s.b, s.h, s.w, s.c, *targs(s))
    Ye = torch.empty(s.P * s.kn, device=DEV)
    cg.gemm("R", "R", "R", "N", "N", s.P, s.kn, s.K, 1.0, rows, s.K, dev(W), s.kn, 0.0, Ye, s.kn)
    Yi = torch.empty(s.P * s.kn, device=DEV)
    cg.sconvGemmNHWC("N", s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil, s.hdil,
                     dev(W), dev(X), Yi)
    assert rel_err(host(Ye), host(Yi))[0] <= 1e-5
    # NCHW: tmp[kn][P] = W[kn][K] cols[K][P], then transpose_nchw to [b][kn][ho][wo]
    X = rng.standard_normal(s.x_shape(NCHW)).astype(np.float32)
    W = rng.standard_normal(s.w_shape(NCHW)).astype(np.float32)
    cols = torch.zeros(s.K * s.P, device=DEV)
    cg.im2col_nchw(cols, s.P, dev(X), s.b, s.c, s.h, s.w, *targs(s))
    tmp = torch.empty(s.kn * s.P, device=DEV)
    cg.gemm("R", "R", "R", "N", "N", s.kn, s.P, s.K, 1.0, dev(W), s.K, cols, s.P, 0.0, tmp, s.P)
    Ye = torch.empty(s.P * s.kn, device=DEV)
    cg.transpose_nchw(s.P, s.kn, tmp, s.P, 0.0, Ye, s.kn, s.ho, s.wo, 0, 0)
    Yi = torch.empty(s.P * s.kn, device=DEV)
    cg.sconvGemmNCHW("N", s.b, s.c, s.h, s.w, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil, s.hdil,
                     dev(W), dev(X), Yi)
    assert rel_err(host(Ye), host(Yi))[0] <= 1e-5


@pytest.mark.parametrize("esize", [4, 2, 1])
@pytest.mark.parametrize("dims", [(2, 5, 7, 3, 3, 2, 4), (3, 37, 3, 3, 1, 1, 6), (1, 300, 4, 4, 3, 3, 2)])
def test_layout_converters(esize, dims):
    """NCHW<->NHWC and weight-order converters; the second and third shapes put
    16 or fewer elements on the inner axis (the short-permute kernel)."""
    dt = {4: torch.float32, 2: torch.int16, 1: torch.uint8}[esize]
    g = torch.Generator().manual_seed(esize)
    b, c, h, w, kh, kw, kn = dims
    x = torch.randint(0, 100, (b, c, h, w), generator=g).to(dt).to(DEV)
    y = torch.empty(b, h, w, c, dtype=dt, device=DEV)
    cg.nchw_to_nhwc(x, y, b, c, h, w, esize)
    torch.cuda.synchronize()
    assert torch.equal(y, x.permute(0, 2, 3, 1).contiguous())
    z = torch.empty_like(x)
    cg.nhwc_to_nchw(y, z, b, h, w, c, esize)
    torch.cuda.synchronize()
    assert torch.equal(z, x)
    wn = torch.randint(0, 100, (c, kh, kw, kn), generator=g).to(dt).to(DEV)  # [c][kh][kw][kn]
    wc = torch.empty(kn, c, kh, kw, dtype=dt, device=DEV)
    cg.weights_nhwc_to_nchw(wn, wc, c, kh, kw, kn, esize)
    torch.cuda.synchronize()
    assert torch.equal(wc, wn.permute(3, 0, 1, 2).contiguous())
    back = torch.empty_like(wn)
    cg.weights_nchw_to_nhwc(wc, back, kn, c, kh, kw, esize)
    torch.cuda.synchronize()
    assert torch.equal(back, wn)
