"""FP64 oracle evaluated at sampled output coordinates (numpy).

TEST INFRASTRUCTURE ONLY.  Used to check full-size problems (e.g. ResNet-50
layers at batch 32) where the direct C oracle would take minutes: the same
formulas as conv_oracle.c (src/convGemm.c:48-146, index maps of
src/im2row_nhwc.c:53-58), evaluated only at chosen output elements.
tests/test_oracle.py pins every function here against the C oracle.
"""
import numpy as np

from . import NCHW, NHWC, Shape

f64 = np.float64


def _to_nhwc_x(X, layout):
    return X if layout == NHWC else X.transpose(0, 2, 3, 1)


def _to_nhwc_w(W, layout):
    # -> [c][kh][kw][kn]
    return W if layout == NHWC else W.transpose(1, 2, 3, 0)


def _to_nhwc_y(Y, layout):
    return Y if layout == NHWC else Y.transpose(0, 2, 3, 1)


def fwd_sampled(layout, s: Shape, W, X, samples, bias=None, mean=None, inv_std=None, gamma=None, bnbeta=None,
                relu=False):
    """Y at samples[i] = (b, x, y, k) with the fused epilogue (alpha=1, beta=0)."""
    Xn, Wn = _to_nhwc_x(X, layout), _to_nhwc_w(W, layout)
    b, x, y, k = (np.asarray(samples)[:, i] for i in range(4))
    ix = x[:, None] * s.vstr + np.arange(s.kh)[None, :] * s.vdil - s.vpad
    iy = y[:, None] * s.hstr + np.arange(s.kw)[None, :] * s.hdil - s.hpad
    vx = (ix >= 0) & (ix < s.h)
    vy = (iy >= 0) & (iy < s.w)
    patch = Xn[b[:, None, None], np.clip(ix, 0, s.h - 1)[:, :, None], np.clip(iy, 0, s.w - 1)[:, None, :], :]
    patch = patch.astype(f64) * (vx[:, :, None] & vy[:, None, :])[..., None]
    w = Wn[:, :, :, k].transpose(3, 1, 2, 0).astype(f64)  # (S, kh, kw, c)
    t = np.einsum("shwc,shwc->s", patch, w)
    if bias is not None:
        t = t + np.asarray(bias, f64)[k]
    if mean is not None:
        t = (t - np.asarray(mean, f64)[k]) * np.asarray(inv_std, f64)[k]
        t = t * np.asarray(gamma, f64)[k] + np.asarray(bnbeta, f64)[k]
    if relu:
        t = np.maximum(t, 0.0)
    return t


def wgrad_sampled(layout, s: Shape, dY, X, samples):
    """dW at samples[i] = (c, kx, ky, k): sum over all output pixels."""
    Xn, Yn = _to_nhwc_x(X, layout), _to_nhwc_y(dY, layout)
    out = np.empty(len(samples))
    xs = np.arange(s.ho)
    ys = np.arange(s.wo)
    for i, (c, kx, ky, k) in enumerate(np.asarray(samples)):
        ix = xs * s.vstr + kx * s.vdil - s.vpad
        iy = ys * s.hstr + ky * s.hdil - s.hpad
        mx = (ix >= 0) & (ix < s.h)
        my = (iy >= 0) & (iy < s.w)
        # (b, nx, ny) planes of channel c / k only (no copy of the other channels)
        xv = Xn[:, ix[mx][:, None], iy[my][None, :], c].astype(f64)
        yv = Yn[:, xs[mx][:, None], ys[my][None, :], k].astype(f64)
        out[i] = np.sum(xv * yv)
    return out


def dgrad_sampled(layout, s: Shape, W, dY, samples):
    """row2im(W^T dY) at samples[i] = (b, ih, iw, c) (without the accumulated dX)."""
    Wn, Yn = _to_nhwc_w(W, layout), _to_nhwc_y(dY, layout)
    out = np.zeros(len(samples))
    for i, (b, ih, iw, c) in enumerate(np.asarray(samples)):
        acc = 0.0
        for kx in range(s.kh):
            num = ih + s.vpad - kx * s.vdil
            if num % s.vstr:
                continue
            x = num // s.vstr
            if x < 0 or x >= s.ho:
                continue
            for ky in range(s.kw):
                numy = iw + s.hpad - ky * s.hdil
                if numy % s.hstr:
                    continue
                y = numy // s.hstr
                if y < 0 or y >= s.wo:
                    continue
                acc += float(np.dot(Yn[b, x, y, :].astype(f64), Wn[c, kx, ky, :].astype(f64)))
        out[i] = acc
    return out


def nhwc_index(layout, arr, idx):
    """Read arr at NHWC-ordered 4-tuples idx (converting for NCHW tensors)."""
    a = _to_nhwc_x(arr, layout)
    idx = np.asarray(idx)
    return a[idx[:, 0], idx[:, 1], idx[:, 2], idx[:, 3]]


def w_index(layout, arr, idx):
    """Read weights at (c, kx, ky, k) tuples."""
    a = _to_nhwc_w(arr, layout)
    idx = np.asarray(idx)
    return a[idx[:, 0], idx[:, 1], idx[:, 2], idx[:, 3]]
