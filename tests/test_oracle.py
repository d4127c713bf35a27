"""Pins the oracle (CPU only).

The FP64 direct-convolution oracle (oracle/conv_oracle.c) is checked against
independent routines and mathematical facts, chosen so that a dropped term,
a wrong sign or index, or a transposed operand fails at least one check:

* torch's float64 CPU conv2d (forward) and its autograd (weight gradient and
  backward data) -- an independent library implementation;
* torch's unfold/fold for the explicit im2col/im2row/col2im/row2im transforms;
* the adjoint identities <fwd(X,W),dY> = <X,dgrad(W,dY)> = <W,wgrad(X,dY)>;
* 1x1 convolution == plain GEMM, full-window convolution == fully-connected;
* the element-wise epilogue closed form and alpha/beta.

The FP32 restatement of the reference algorithm (oracle/convgemm_port.c) is then
checked against the pinned FP64 oracle, including multi-block K (K > KC = 256),
multi-block N (n > NC = 3072) and the reference's own acceptance test
(element-wise relative 1e-5 on U[0,1] data, tests/test_base.h:19-29).
"""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

import oracle
from oracle import NCHW, NHWC, Shape

RNG = np.random.default_rng(0xC0FFEE)

SHAPES = [
    Shape(2, 7, 6, 3, 4, 3, 3, 1, 1, 1, 1, 1, 1),
    Shape(1, 9, 8, 5, 6, 3, 2, 1, 0, 2, 1, 1, 1),
    Shape(2, 10, 11, 4, 3, 3, 3, 2, 1, 1, 2, 2, 1),   # dilation, asymmetric stride/pad
    Shape(1, 8, 8, 6, 5, 1, 1, 0, 0, 1, 1, 1, 1),     # 1x1
    Shape(3, 5, 5, 2, 7, 5, 5, 2, 2, 2, 2, 1, 1),     # 5x5 s2
    Shape(1, 13, 12, 3, 2, 4, 3, 0, 3, 3, 2, 1, 2),   # even kernel, stride 3
    Shape(2, 4, 4, 3, 3, 4, 4, 0, 0, 1, 1, 1, 1),     # full window
]


def torch_nchw(x, layout):
    t = torch.from_numpy(np.ascontiguousarray(x))
    return t if layout == NCHW else t.permute(0, 3, 1, 2)


def torch_w(w, layout):
    t = torch.from_numpy(np.ascontiguousarray(w))
    return t if layout == NCHW else t.permute(3, 0, 1, 2)  # [c][kh][kw][kn] -> [kn][c][kh][kw]


def from_nchw(t, layout):
    t = t.detach()
    return (t if layout == NCHW else t.permute(0, 2, 3, 1)).contiguous().numpy()


def w_from_torch(t, layout):
    t = t.detach()
    return (t if layout == NCHW else t.permute(1, 2, 3, 0)).contiguous().numpy()


def torch_conv(s, X, W, layout):
    return F.conv2d(torch_nchw(X, layout), torch_w(W, layout), stride=(s.vstr, s.hstr),
                    padding=(s.vpad, s.hpad), dilation=(s.vdil, s.hdil))


@pytest.mark.parametrize("layout", [NHWC, NCHW])
@pytest.mark.parametrize("si", range(len(SHAPES)))
def test_fwd_matches_torch_conv2d(layout, si):
    s = SHAPES[si]
    X = RNG.standard_normal(s.x_shape(layout))
    W = RNG.standard_normal(s.w_shape(layout))
    Y = oracle.fwd(layout, s, W, X)
    ref = from_nchw(torch_conv(s, X, W, layout), layout)
    assert Y.shape == ref.shape
    np.testing.assert_allclose(Y, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("layout", [NHWC, NCHW])
@pytest.mark.parametrize("si", range(len(SHAPES)))
def test_wgrad_dgrad_match_torch_autograd(layout, si):
    s = SHAPES[si]
    X = RNG.standard_normal(s.x_shape(layout))
    W = RNG.standard_normal(s.w_shape(layout))
    dY = RNG.standard_normal(s.y_shape(layout))
    tx = torch_nchw(X, layout).clone().requires_grad_(True)
    tw = torch_w(W, layout).clone().requires_grad_(True)
    y = F.conv2d(tx, tw, stride=(s.vstr, s.hstr), padding=(s.vpad, s.hpad), dilation=(s.vdil, s.hdil))
    gx, gw = torch.autograd.grad((y * torch_nchw(dY, layout)).sum(), (tx, tw))
    np.testing.assert_allclose(oracle.wgrad(layout, s, dY, X), w_from_torch(gw, layout), rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(oracle.dgrad(layout, s, W, dY), from_nchw(gx, layout), rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("layout", [NHWC, NCHW])
@pytest.mark.parametrize("si", range(len(SHAPES)))
def test_adjoint_identities(layout, si):
    s = SHAPES[si]
    X = RNG.standard_normal(s.x_shape(layout))
    W = RNG.standard_normal(s.w_shape(layout))
    dY = RNG.standard_normal(s.y_shape(layout))
    a = np.vdot(oracle.fwd(layout, s, W, X), dY)
    b = np.vdot(X, oracle.dgrad(layout, s, W, dY))
    c = np.vdot(W, oracle.wgrad(layout, s, dY, X))
    scale = max(1.0, abs(a))
    assert abs(a - b) <= 1e-12 * scale * s.K
    assert abs(a - c) <= 1e-12 * scale * s.K


@pytest.mark.parametrize("layout", [NHWC, NCHW])
def test_1x1_is_gemm(layout):
    s = Shape(2, 5, 4, 7, 3, 1, 1)
    X = RNG.standard_normal(s.x_shape(layout))
    W = RNG.standard_normal(s.w_shape(layout))
    Y = oracle.fwd(layout, s, W, X)
    if layout == NHWC:
        ref = X.reshape(-1, s.c) @ W.reshape(s.c, s.kn)
        np.testing.assert_allclose(Y.reshape(-1, s.kn), ref, rtol=1e-13, atol=1e-13)
    else:
        ref = np.einsum("kc,bcp->bkp", W.reshape(s.kn, s.c), X.reshape(s.b, s.c, -1))
        np.testing.assert_allclose(Y.reshape(s.b, s.kn, -1), ref, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("layout", [NHWC, NCHW])
def test_full_window_is_fully_connected(layout):
    s = Shape(3, 4, 5, 2, 6, 4, 5)
    X = RNG.standard_normal(s.x_shape(layout))
    W = RNG.standard_normal(s.w_shape(layout))
    Y = oracle.fwd(layout, s, W, X)
    assert s.ho == 1 and s.wo == 1
    if layout == NHWC:
        # r = (c*kh + kx)*kw + ky: X[b][kx][ky][c] reordered to [c][kx][ky]
        Xf = X.transpose(0, 3, 1, 2).reshape(s.b, -1)
        ref = Xf @ W.reshape(s.K, s.kn)
    else:
        ref = X.reshape(s.b, -1) @ W.reshape(s.kn, s.K).T
    np.testing.assert_allclose(Y.reshape(s.b, s.kn), ref, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("layout", [NHWC, NCHW])
@pytest.mark.parametrize("flags", [(1, 1, 1), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (0, 1, 1), (1, 0, 1)])
def test_epilogue_closed_form(layout, flags):
    """1x1 identity weights make Y = epi(X); compare with the formula of
    im2row_nhwc.c:276-282 written element-wise."""
    use_bias, use_bn, relu = flags
    s = Shape(2, 3, 4, 5, 5, 1, 1)
    X = RNG.standard_normal(s.x_shape(layout))
    W = np.eye(5).reshape(s.w_shape(layout)) if layout == NHWC else np.eye(5).reshape(5, 5, 1, 1)
    Y0 = RNG.standard_normal(s.y_shape(layout))
    bias = RNG.standard_normal(5) if use_bias else None
    mean, inv_std, gamma, bnb = (RNG.standard_normal(5), RNG.uniform(0.5, 2, 5), RNG.uniform(0.5, 1.5, 5),
                                 RNG.standard_normal(5)) if use_bn else (None,) * 4
    alpha, beta = 0.75, 1.0
    Y = oracle.fwd(layout, s, W, X, alpha=alpha, beta=beta, Y0=Y0, bias=bias, mean=mean, inv_std=inv_std,
                   gamma=gamma, bnbeta=bnb, relu=bool(relu))
    ch_axis = 3 if layout == NHWC else 1
    shp = [1, 1, 1, 1]
    shp[ch_axis] = 5
    t = alpha * X + beta * Y0
    if use_bias:
        t = t + bias.reshape(shp)
    if use_bn:
        t = (t - mean.reshape(shp)) * inv_std.reshape(shp) * gamma.reshape(shp) + bnb.reshape(shp)
    if relu:
        t = np.maximum(t, 0.0)
    np.testing.assert_allclose(Y, t, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("layout", [NHWC, NCHW])
def test_alpha_beta_linearity(layout):
    s = SHAPES[2]
    X = RNG.standard_normal(s.x_shape(layout))
    W = RNG.standard_normal(s.w_shape(layout))
    dY = RNG.standard_normal(s.y_shape(layout))
    base_y = oracle.fwd(layout, s, W, X)
    Y0 = RNG.standard_normal(base_y.shape)
    np.testing.assert_allclose(oracle.fwd(layout, s, W, X, alpha=-2.5, beta=0.5, Y0=Y0), -2.5 * base_y + 0.5 * Y0,
                               rtol=1e-12, atol=1e-12)
    base_w = oracle.wgrad(layout, s, dY, X)
    W0 = RNG.standard_normal(base_w.shape)
    np.testing.assert_allclose(oracle.wgrad(layout, s, dY, X, alpha=3.0, beta=-1.5, dW0=W0), 3 * base_w - 1.5 * W0,
                               rtol=1e-12, atol=1e-12)
    base_x = oracle.dgrad(layout, s, W, dY, beta=0.0)
    X0 = RNG.standard_normal(base_x.shape)
    np.testing.assert_allclose(oracle.dgrad(layout, s, W, dY, alpha=0.5, beta=2.0, dX0=X0), 0.5 * base_x + 2 * X0,
                               rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("si", range(len(SHAPES)))
def test_explicit_transforms_match_unfold_fold(si):
    s = SHAPES[si]
    kw = dict(kernel_size=(s.kh, s.kw), dilation=(s.vdil, s.hdil), padding=(s.vpad, s.hpad), stride=(s.vstr, s.hstr))
    Xc = RNG.standard_normal(s.x_shape(NCHW))
    U = F.unfold(torch.from_numpy(Xc), **kw).numpy()  # [b][K][L]
    cols = oracle.im2col_nchw(s, Xc)  # [K][P], P = (b, L)
    np.testing.assert_array_equal(cols, U.transpose(1, 0, 2).reshape(s.K, -1))
    Xh = Xc.transpose(0, 2, 3, 1).copy()
    rows = oracle.im2row_nhwc(s, Xh)  # [P][K]
    np.testing.assert_array_equal(rows, U.transpose(0, 2, 1).reshape(-1, s.K))
    # col2im / row2im are the adjoint scatters: fold
    C = RNG.standard_normal((s.K, s.P))
    fold = F.fold(torch.from_numpy(C.reshape(s.K, s.b, -1).transpose(1, 0, 2).copy()), output_size=(s.h, s.w),
                  **kw).numpy()
    np.testing.assert_allclose(oracle.col2im_nchw(s, C), fold, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(oracle.row2im_nhwc(s, C.T.copy()), fold.transpose(0, 2, 3, 1), rtol=1e-13,
                               atol=1e-13)


def test_swap01():
    a = RNG.standard_normal((3, 4, 2, 5))
    np.testing.assert_array_equal(oracle.swap01(a), a.transpose(1, 0, 2, 3))
    o0 = RNG.standard_normal((4, 3, 2, 5))
    np.testing.assert_allclose(oracle.swap01(a, beta=2.0, out0=o0), 2 * o0 + a.transpose(1, 0, 2, 3))


def test_out_size_c_division():
    assert oracle.out_size(56, 1, 1, 3, 1) == 56
    assert oracle.out_size(56, 1, 1, 3, 2) == 28
    assert oracle.out_size(55, 0, 1, 11, 4) == 12
    assert oracle.out_size(2, 0, 1, 5, 1) == -2  # (2-4-1)/1 + 1
    assert oracle.out_size(2, 0, 1, 5, 2) == 0  # (-3)/2 = -1 in C (truncation), +1 = 0


# ---------------------------------------------------------------- FP32 port


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


PORT_SHAPES = [
    Shape(2, 9, 7, 5, 6, 3, 3, 1, 1, 1, 1, 1, 1),
    Shape(1, 11, 10, 4, 7, 3, 2, 2, 0, 2, 1, 1, 2),
    Shape(2, 8, 8, 40, 20, 3, 3, 1, 1, 1, 1, 1, 1),   # K = 360 > KC = 256 (two K blocks)
    Shape(4, 30, 30, 3, 5, 1, 1, 0, 0, 1, 1, 1, 1),   # P = 3600 > NC = 3072 (two N blocks)
    Shape(1, 6, 6, 3, 2, 7, 7, 3, 3, 1, 1, 1, 1),
]


@pytest.mark.parametrize("layout", [NHWC, NCHW])
@pytest.mark.parametrize("si", range(len(PORT_SHAPES)))
def test_port_forward_with_epilogue(layout, si):
    s = PORT_SHAPES[si]
    X = RNG.standard_normal(s.x_shape(layout)).astype(np.float32)
    W = (RNG.standard_normal(s.w_shape(layout)) / np.sqrt(s.K)).astype(np.float32)
    epi = dict(bias=RNG.standard_normal(s.kn).astype(np.float32),
               mean=RNG.standard_normal(s.kn).astype(np.float32),
               inv_std=RNG.uniform(0.5, 2, s.kn).astype(np.float32),
               gamma=RNG.uniform(0.5, 1.5, s.kn).astype(np.float32),
               bnbeta=RNG.standard_normal(s.kn).astype(np.float32), relu=True)
    Y = oracle.port_conv(layout, "N", s, W, X, **epi)
    ref = oracle.fwd(layout, s, W, X, **epi)
    assert _rel(Y, ref) < 1e-5
    Y2 = oracle.port_conv(layout, "N", s, W, X)  # no epilogue
    assert _rel(Y2, oracle.fwd(layout, s, W, X)) < 1e-5


@pytest.mark.parametrize("layout", [NHWC, NCHW])
@pytest.mark.parametrize("si", range(len(PORT_SHAPES)))
def test_port_wgrad_dgrad(layout, si):
    s = PORT_SHAPES[si]
    X = RNG.standard_normal(s.x_shape(layout)).astype(np.float32)
    W = RNG.standard_normal(s.w_shape(layout)).astype(np.float32)
    dY = RNG.standard_normal(s.y_shape(layout)).astype(np.float32)
    dW = oracle.port_conv(layout, "T", s, dY, X)
    assert _rel(dW, oracle.wgrad(layout, s, dY, X)) < 1e-5
    dX0 = RNG.standard_normal(s.x_shape(layout)).astype(np.float32)
    dX = oracle.port_back(layout, s, W, dY, dX0.copy())
    assert _rel(dX, oracle.dgrad(layout, s, W, dY, dX0=dX0)) < 1e-5  # accumulates: dx += ...


@pytest.mark.parametrize("layout", [NHWC, NCHW])
def test_port_reference_acceptance_on_uniform_data(layout):
    """The reference's own check: element-wise |a-b|/|a| <= 1e-5 on U[0,1] data
    with bias+BN+ReLU all on (tests/test_base.c:13-15, tests/test_base.h:12-29)."""
    s = Shape(2, 12, 12, 16, 24, 3, 3, 1, 1, 1, 1, 1, 1)
    u = lambda *shape: RNG.uniform(0, 1, shape).astype(np.float32)
    X, W = u(*s.x_shape(layout)), u(*s.w_shape(layout))
    epi = dict(bias=u(s.kn), mean=u(s.kn), inv_std=u(s.kn), gamma=u(s.kn), bnbeta=u(s.kn), relu=True)
    Y = oracle.port_conv(layout, "N", s, W, X, **epi).astype(np.float64)
    ref = oracle.fwd(layout, s, W, X, **epi)
    nz = ref != 0
    assert np.all(np.abs((ref[nz] - Y[nz]) / ref[nz]) <= 1e-5)


# ---------------------------------------------------------------- sampled oracle

@pytest.mark.parametrize("layout", [NHWC, NCHW])
@pytest.mark.parametrize("si", range(len(SHAPES)))
def test_sampled_oracle_matches_direct(layout, si):
    from oracle import sampled
    s = SHAPES[si]
    X = RNG.standard_normal(s.x_shape(layout))
    W = RNG.standard_normal(s.w_shape(layout))
    dY = RNG.standard_normal(s.y_shape(layout))
    epi = dict(bias=RNG.standard_normal(s.kn), mean=RNG.standard_normal(s.kn), inv_std=RNG.uniform(0.5, 2, s.kn),
               gamma=RNG.uniform(0.5, 1.5, s.kn), bnbeta=RNG.standard_normal(s.kn), relu=True)
    Y = oracle.fwd(layout, s, W, X, **epi)
    ys = np.stack([RNG.integers(0, n, 40) for n in (s.b, s.ho, s.wo, s.kn)], 1)
    np.testing.assert_allclose(sampled.fwd_sampled(layout, s, W, X, ys, **epi), sampled.nhwc_index(layout, Y, ys),
                               rtol=1e-12, atol=1e-12)
    dW = oracle.wgrad(layout, s, dY, X)
    ws = np.stack([RNG.integers(0, n, 40) for n in (s.c, s.kh, s.kw, s.kn)], 1)
    np.testing.assert_allclose(sampled.wgrad_sampled(layout, s, dY, X, ws), sampled.w_index(layout, dW, ws),
                               rtol=1e-12, atol=1e-12)
    dX = oracle.dgrad(layout, s, W, dY, beta=0.0)
    xs = np.stack([RNG.integers(0, n, 40) for n in (s.b, s.h, s.w, s.c)], 1)
    np.testing.assert_allclose(sampled.dgrad_sampled(layout, s, W, dY, xs), sampled.nhwc_index(layout, dX, xs),
                               rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------- MLP oracle

def test_mlp_oracle_gradient_matches_finite_differences():
    from oracle import mlp
    rng = np.random.default_rng(12)
    p = rng.standard_normal(mlp.NP) * 0.2
    X = rng.standard_normal((17, mlp.F))
    T = rng.standard_normal(17)
    L, g = mlp.loss_grad(p, X, T)
    assert abs(L - mlp.loss(p, X, T)) < 1e-14
    idx = rng.choice(mlp.NP, 200, replace=False)
    idx = np.concatenate([idx, [mlp.NP - 1, mlp.F * mlp.H, 0]])  # include b3, b1[0], W1[0,0]
    h = 1e-6
    for i in idx:
        e = np.zeros(mlp.NP)
        e[i] = h
        fd = (mlp.loss(p + e, X, T) - mlp.loss(p - e, X, T)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_mlp_oracle_matches_torch_autograd():
    """The MLP oracle's forward pass and loss gradient against an independent
    implementation: torch's float64 linear layers and autograd on the same
    flat parameter layout [W1 (FxH), b1, W2 (HxH), b2, W3 (Hx1), b3]."""
    import torch
    from oracle import mlp
    rng = np.random.default_rng(12)
    p = rng.standard_normal(mlp.NP) * 0.2
    X = rng.standard_normal((23, mlp.F))
    T = rng.standard_normal(23)
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    F, H = mlp.F, mlp.H
    o = [0, F * H, F * H + H, F * H + H + H * H, F * H + 2 * H + H * H, F * H + 3 * H + H * H, mlp.NP]
    W1, b1, W2, b2, W3, b3 = (pt[o[i]:o[i + 1]] for i in range(6))
    Xt = torch.tensor(X, dtype=torch.float64)
    h1 = torch.relu(torch.nn.functional.linear(Xt, W1.reshape(F, H).t(), b1))
    h2 = torch.relu(torch.nn.functional.linear(h1, W2.reshape(H, H).t(), b2))
    y = torch.nn.functional.linear(h2, W3.reshape(H, 1).t(), b3)[:, 0]
    loss = torch.mean((y - torch.tensor(T, dtype=torch.float64)) ** 2)
    loss.backward()
    y_o, _ = mlp.forward(p, X)
    L_o, g_o = mlp.loss_grad(p, X, T)
    np.testing.assert_allclose(y_o, y.detach().numpy(), rtol=1e-12, atol=1e-12)
    assert abs(L_o - loss.item()) <= 1e-12 * max(1.0, abs(L_o))
    np.testing.assert_allclose(g_o, pt.grad.numpy(), rtol=1e-10, atol=1e-12)
