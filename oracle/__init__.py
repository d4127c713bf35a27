"""Oracle for the convolution-as-implicit-GEMM path.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py`` -- always as the checker, never
as the thing measured or shipped.  The product library (``convgemm_amd``)
never imports this package.

Two C libraries are wrapped here:

* ``liboracle.so`` (conv_oracle.c): FP64 direct convolution for the six
  operations of the reference (forward / weight-gradient / backward-data, NHWC
  and NCHW, src/convGemm.c:48-146), the fused bias/BN/ReLU epilogue
  (src/im2row_nhwc.c:269-372, src/im2col_nchw.c:397-493) and the explicit
  transforms im2row/row2im/im2col/col2im/transpose (src/im2row_nhwc.c:147-229,
  src/im2col_nchw.c:146-233, :337-359).
* ``libconvgemm_port.so`` (convgemm_port.c): FP32 restatement of the
  reference's BLIS-style five-loop implicit GEMM (src/gemm_blis_B3A2C0*.c),
  with the same four entry points as src/convGemm.h:22-64 (prefixed
  ``port_``).  It is the CPU baseline ("port") and a second FP32 checker.

Pinning (see tests/test_oracle.py): the FP64 functions are checked against
torch's float64 CPU ``conv2d``/autograd and ``unfold``/``fold`` (independent
library routines), the adjoint identities <fwd(X,W),dY> = <X,dgrad> =
<W,wgrad>, 1x1 == plain GEMM, full-window == fully-connected layer, and the
element-wise epilogue closed form.  The FP32 port is checked against the
FP64 functions.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

NHWC, NCHW = 0, 1


def build(force: bool = False) -> None:
    """Compile the oracle's C sources in place (gcc + OpenMP)."""
    targets = ["liboracle.so", "libconvgemm_port.so"]
    if force or not all(os.path.exists(os.path.join(_HERE, t)) for t in targets):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)


def _lib(name: str) -> ctypes.CDLL:
    if name not in _LIBS:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            build()
        _LIBS[name] = ctypes.CDLL(path)
    return _LIBS[name]


_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)


def _d(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _f(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_fp)


def _shape_arr(s):
    return (ctypes.c_int * 13)(*[int(v) for v in s])


def out_size(inp, pad, dil, k, stride):
    """ho = (h + 2*pad - dil*(k-1) - 1)/stride + 1, C division (src/convGemm.c:60-61)."""
    num = inp + 2 * pad - dil * (k - 1) - 1
    q = abs(num) // stride
    q = q if num >= 0 else -q
    return q + 1


class Shape:
    """Convolution problem: b,h,w,c,kn,kh,kw,vpad,hpad,vstr,hstr,vdil,hdil."""

    FIELDS = ("b", "h", "w", "c", "kn", "kh", "kw", "vpad", "hpad", "vstr", "hstr", "vdil", "hdil")

    def __init__(self, b, h, w, c, kn, kh, kw, vpad=0, hpad=0, vstr=1, hstr=1, vdil=1, hdil=1):
        self.b, self.h, self.w, self.c, self.kn, self.kh, self.kw = b, h, w, c, kn, kh, kw
        self.vpad, self.hpad, self.vstr, self.hstr, self.vdil, self.hdil = vpad, hpad, vstr, hstr, vdil, hdil

    @property
    def ho(self):
        return out_size(self.h, self.vpad, self.vdil, self.kh, self.vstr)

    @property
    def wo(self):
        return out_size(self.w, self.hpad, self.hdil, self.kw, self.hstr)

    @property
    def K(self):
        return self.c * self.kh * self.kw

    @property
    def P(self):
        return self.b * self.ho * self.wo

    def tuple(self):
        return tuple(getattr(self, f) for f in self.FIELDS)

    def flops(self):
        """2*P*K*N, padding zeros included (plots/convgemm.plt:56-57)."""
        return 2.0 * self.P * self.K * self.kn

    def x_shape(self, layout):
        return (self.b, self.h, self.w, self.c) if layout == NHWC else (self.b, self.c, self.h, self.w)

    def w_shape(self, layout):
        return (self.c, self.kh, self.kw, self.kn) if layout == NHWC else (self.kn, self.c, self.kh, self.kw)

    def y_shape(self, layout):
        return (self.b, self.ho, self.wo, self.kn) if layout == NHWC else (self.b, self.kn, self.ho, self.wo)

    def __repr__(self):
        return "Shape(" + ", ".join(f"{f}={getattr(self, f)}" for f in self.FIELDS) + ")"


def _c64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def fwd(layout, s: Shape, W, X, alpha=1.0, beta=0.0, Y0=None, bias=None, mean=None,
        inv_std=None, gamma=None, bnbeta=None, relu=False):
    """FP64 forward: Y = epi(alpha*conv(X,W) + beta*Y0)."""
    lib = _lib("liboracle.so")
    W, X = _c64(W), _c64(X)
    Y = np.zeros(s.y_shape(layout)) if Y0 is None else _c64(Y0).copy()
    bias, mean, inv_std, gamma, bnbeta = map(_c64, (bias, mean, inv_std, gamma, bnbeta))
    lib.oracle_fwd(ctypes.c_int(layout), _shape_arr(s.tuple()), ctypes.c_double(alpha),
                   ctypes.c_double(beta), _d(W), _d(X), _d(Y), _d(bias), _d(mean), _d(inv_std),
                   _d(gamma), _d(bnbeta), ctypes.c_int(int(bool(relu))))
    return Y


def wgrad(layout, s: Shape, dY, X, alpha=1.0, beta=0.0, dW0=None):
    """FP64 weight gradient: dW = alpha*im2row(X)^T dY + beta*dW0 (dW in W's layout)."""
    lib = _lib("liboracle.so")
    dY, X = _c64(dY), _c64(X)
    dW = np.zeros(s.w_shape(layout)) if dW0 is None else _c64(dW0).copy()
    lib.oracle_wgrad(ctypes.c_int(layout), _shape_arr(s.tuple()), ctypes.c_double(alpha),
                     ctypes.c_double(beta), _d(dY), _d(X), _d(dW))
    return dW


def dgrad(layout, s: Shape, W, dY, alpha=1.0, beta=1.0, dX0=None):
    """FP64 backward data: dX = alpha*row2im(W^T dY) + beta*dX0 (the API: alpha=beta=1)."""
    lib = _lib("liboracle.so")
    W, dY = _c64(W), _c64(dY)
    dX = np.zeros(s.x_shape(layout)) if dX0 is None else _c64(dX0).copy()
    lib.oracle_dgrad(ctypes.c_int(layout), _shape_arr(s.tuple()), ctypes.c_double(alpha),
                     ctypes.c_double(beta), _d(W), _d(dY), _d(dX))
    return dX


def im2row_nhwc(s: Shape, X):
    lib = _lib("liboracle.so")
    rows = np.zeros((s.P, s.K))
    lib.oracle_im2row_nhwc(_shape_arr(s.tuple()), _d(_c64(X)), _d(rows))
    return rows


def row2im_nhwc(s: Shape, rows, X0=None):
    lib = _lib("liboracle.so")
    X = np.zeros(s.x_shape(NHWC)) if X0 is None else _c64(X0).copy()
    lib.oracle_row2im_nhwc(_shape_arr(s.tuple()), _d(_c64(rows)), _d(X))
    return X


def im2col_nchw(s: Shape, X):
    lib = _lib("liboracle.so")
    cols = np.zeros((s.K, s.P))
    lib.oracle_im2col_nchw(_shape_arr(s.tuple()), _d(_c64(X)), _d(cols))
    return cols


def col2im_nchw(s: Shape, cols, X0=None):
    lib = _lib("liboracle.so")
    X = np.zeros(s.x_shape(NCHW)) if X0 is None else _c64(X0).copy()
    lib.oracle_col2im_nchw(_shape_arr(s.tuple()), _d(_c64(cols)), _d(X))
    return X


def swap01(arr, beta=0.0, out0=None):
    """out[j][i][...] = in[i][j][...] (+ beta*out0): transpose_nchw's operation."""
    lib = _lib("liboracle.so")
    a = _c64(arr)
    d0, d1 = a.shape[0], a.shape[1]
    inner = int(np.prod(a.shape[2:])) if a.ndim > 2 else 1
    out = np.zeros((d1, d0) + a.shape[2:]) if out0 is None else _c64(out0).copy()
    lib.oracle_swap01(ctypes.c_int(d0), ctypes.c_int(d1), ctypes.c_int(inner), ctypes.c_double(beta),
                      _d(a), _d(out))
    return out


# ---------------------------------------------------------------- FP32 port

def port_set_threads(n: int) -> None:
    _lib("libconvgemm_port.so").port_set_threads(ctypes.c_int(int(n)))


def port_get_threads() -> int:
    return int(_lib("libconvgemm_port.so").port_get_threads())


def _c32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def _ints(*v):
    return [ctypes.c_int(int(x)) for x in v]


def port_conv(layout, trans, s: Shape, inp, x, out=None, bias=None, mean=None, inv_std=None,
              gamma=None, bnbeta=None, relu=False):
    """FP32 restatement of sconvGemmNHWC / sconvGemmNCHW (forward or weight gradient)."""
    lib = _lib("libconvgemm_port.so")
    inp, x = _c32(inp), _c32(x)
    bias, mean, inv_std, gamma, bnbeta = map(_c32, (bias, mean, inv_std, gamma, bnbeta))
    if out is None:
        out = np.zeros(s.y_shape(layout) if trans == "N" else s.w_shape(layout), dtype=np.float32)
    assert out.dtype == np.float32 and out.flags.c_contiguous
    if layout == NHWC:
        dims = _ints(s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil, s.hdil)
        fn = lib.port_sconvGemmNHWC
    else:
        dims = _ints(s.b, s.c, s.h, s.w, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil, s.hdil)
        fn = lib.port_sconvGemmNCHW
    fn(ctypes.c_char(trans.encode()), *dims, _f(inp), _f(x), _f(out), _f(bias), _f(mean), _f(inv_std),
       _f(gamma), _f(bnbeta), ctypes.c_bool(bool(relu)))
    return out


def port_back(layout, s: Shape, weights, dy, dx):
    """FP32 restatement of sconvGemmNHWC_back / sconvGemmNCHW_back (dx += ...)."""
    lib = _lib("libconvgemm_port.so")
    weights, dy = _c32(weights), _c32(dy)
    assert dx.dtype == np.float32 and dx.flags.c_contiguous
    if layout == NHWC:
        dims = _ints(s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vstr, s.hstr, s.vpad, s.hpad, s.vdil, s.hdil)
        fn = lib.port_sconvGemmNHWC_back
    else:
        dims = _ints(s.b, s.c, s.h, s.w, s.kn, s.kh, s.kw, s.vstr, s.hstr, s.vpad, s.hpad, s.vdil, s.hdil)
        fn = lib.port_sconvGemmNCHW_back
    fn(*dims, _f(weights), _f(dy), _f(dx))
    return dx
