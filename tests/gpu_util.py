"""Helpers for the GPU parity tests (imported by tests marked `gpu`)."""
import numpy as np
import torch

import oracle
from oracle import NCHW, NHWC, Shape

import convgemm_amd as cg

DEV = "cuda:0"


def rel_err(got, ref):
    """max|got - ref| / max|ref| and ||got - ref||_2 / ||ref||_2 (float64)."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    d = np.abs(got - ref)
    mx = max(float(np.max(np.abs(ref))) if ref.size else 0.0, 1e-30)
    nr = max(float(np.linalg.norm(ref)), 1e-30)
    return (float(np.max(d)) / mx if d.size else 0.0), float(np.linalg.norm(got - ref)) / nr


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(DEV)


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def shape_args(s: Shape, layout):
    if layout == NHWC:
        return (s.b, s.h, s.w, s.c)
    return (s.b, s.c, s.h, s.w)


def run_fwd(layout, s, W, X, epi=None, packs=(None, None)):
    """Drop-in forward through the C ABI; returns Y on the host."""
    epi = epi or {}
    Y = torch.full(s.y_shape(layout), float("nan"), device=DEV)  # overwritten (beta = 0)
    e = {k: (dev(v) if v is not None else None) for k, v in epi.items() if k != "relu"}
    fn = cg.sconvGemmNHWC if layout == NHWC else cg.sconvGemmNCHW
    fn("N", *shape_args(s, layout), s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil, s.hdil,
       dev(W), dev(X), Y, e.get("bias"), e.get("mean"), e.get("inv_std"), e.get("gamma"), e.get("bnbeta"),
       bool(epi.get("relu", False)), packs[0], packs[1])
    return host(Y)


def run_wgrad(layout, s, dY, X, packs=(None, None)):
    dW = torch.full(s.w_shape(layout), float("nan"), device=DEV)
    fn = cg.sconvGemmNHWC if layout == NHWC else cg.sconvGemmNCHW
    fn("T", *shape_args(s, layout), s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil, s.hdil,
       dev(dY), dev(X), dW, None, None, None, None, None, False, packs[0], packs[1])
    return host(dW)


def run_dgrad(layout, s, W, dY, dX0, packs=(None, None)):
    dX = dev(dX0)
    fn = cg.sconvGemmNHWC_back if layout == NHWC else cg.sconvGemmNCHW_back
    fn(*shape_args(s, layout), s.kn, s.kh, s.kw, s.vstr, s.hstr, s.vpad, s.hpad, s.vdil, s.hdil,
       dev(W), dev(dY), dX, packs[0], packs[1])
    return host(dX)


def random_problem(rng, s, layout, epi=True):
    X = rng.standard_normal(s.x_shape(layout)).astype(np.float32)
    W = (rng.standard_normal(s.w_shape(layout)) * np.sqrt(2.0 / max(s.K, 1))).astype(np.float32)
    dY = rng.standard_normal(s.y_shape(layout)).astype(np.float32)
    e = None
    if epi:
        e = dict(bias=(0.1 * rng.standard_normal(s.kn)).astype(np.float32),
                 mean=(0.1 * rng.standard_normal(s.kn)).astype(np.float32),
                 inv_std=(1.0 / np.sqrt(rng.uniform(0.5, 2.0, s.kn) + 1e-5)).astype(np.float32),
                 gamma=rng.uniform(0.5, 1.5, s.kn).astype(np.float32),
                 bnbeta=(0.1 * rng.standard_normal(s.kn)).astype(np.float32), relu=True)
    return X, W, dY, e
