"""GPU test of the allocation-free entry point convgemm_conv_noalloc.

The call runs convgemm_conv_ex's kernels on the library's static device
scratch.  Each case makes two calls in a row on the same scratch (the second
with a smaller input) and checks every output against convgemm_conv_ex with
the same plan (bit-identical: the same kernels, the same summation order)
and against the FP64 oracle (FP32 bar 1e-5, max-normwise).  Plans are forced
to ones that use the scratch for both of its parts: repacked weights and
split-K partials.
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
    cg.lib()
    yield
    cg.force_config(-1, 0)


from tests.gpu_util import DEV, dev, host, random_problem, rel_err  # noqa: E402
import convgemm_amd as cg  # noqa: E402


def _desc(op, layout, s):
    return cg.make_desc(op, layout, s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil,
                        s.hdil)


def _general_cfg(op, layout):
    """A general-kernel 64x64 FP32 configuration that serves (op, layout) with 16-byte loads."""
    opk = {(cg.FWD, NHWC): 0, (cg.FWD, NCHW): 1, (cg.WGRAD, NHWC): 2, (cg.WGRAD, NCHW): 3,
           (cg.DGRAD, NHWC): 4, (cg.DGRAD, NCHW): 5}[(op, layout)]
    for i, t in enumerate(cg.tile_configs(cg.F32)):
        if t[0] == 64 and t[1] == 64 and t[5] == 0 and cg.lib().cg_tilecfg_supports(cg.F32, i, opk, 1):
            return i
    raise AssertionError("no 64x64 general configuration")


def _operands(op, layout, s, rng):
    X, W, dY, e = random_problem(rng, s, layout, epi=op == cg.FWD)
    if op == cg.FWD:
        return dev(W), dev(X), (s.y_shape(layout)), e, oracle.fwd(layout, s, W, X, **(e or {}))
    if op == cg.WGRAD:
        return dev(dY), dev(X), (s.w_shape(layout)), None, oracle.wgrad(layout, s, dY, X)
    dX0 = rng.standard_normal(s.x_shape(layout)).astype(np.float32)
    return dev(W), dev(dY), dX0, None, oracle.dgrad(layout, s, W, dY, dX0=dX0)


def _run(fn, op, layout, s, ins, out0, epi, **kw):
    in_, x = ins
    if op == cg.DGRAD:
        out = dev(out0)
    else:
        out = torch.full(out0, float("nan"), device=DEV)
    e = {k: dev(v) for k, v in (epi or {}).items() if k != "relu"}
    fn(_desc(op, layout, s), in_, x, out, beta=1.0 if op == cg.DGRAD else 0.0, relu=bool((epi or {}).get("relu")),
       bias=e.get("bias"), mean=e.get("mean"), inv_std=e.get("inv_std"), gamma=e.get("gamma"),
       bnbeta=e.get("bnbeta"), **kw)
    return host(out)


CASES = [
    # (op, layout, first shape, second (smaller) shape, split-K factor)
    (cg.FWD, NHWC, Shape(4, 28, 28, 64, 96, 3, 3, 1, 1, 1, 1, 1, 1), Shape(2, 14, 14, 32, 48, 3, 3, 1, 1, 1, 1, 1, 1),
     3),
    (cg.FWD, NCHW, Shape(2, 20, 20, 48, 64, 3, 3, 1, 1, 2, 2, 1, 1), Shape(1, 9, 9, 16, 32, 3, 3, 1, 1, 1, 1, 1, 1), 2),
    (cg.WGRAD, NHWC, Shape(8, 28, 28, 64, 64, 3, 3, 1, 1, 1, 1, 1, 1), Shape(2, 14, 14, 32, 32, 3, 3, 1, 1, 1, 1, 1, 1),
     8),
    (cg.DGRAD, NCHW, Shape(4, 16, 16, 64, 32, 3, 3, 1, 1, 2, 2, 1, 1), Shape(2, 8, 8, 32, 16, 3, 3, 1, 1, 1, 1, 1, 1),
     2),
]


@pytest.mark.parametrize("op,layout,s1,s2,splits", CASES)
def test_noalloc_matches_conv_ex(op, layout, s1, s2, splits):
    rng = np.random.default_rng(11 + op + 10 * layout)
    cfg = _general_cfg(op, layout)
    for s in (s1, s2):  # two calls in a row on the shared scratch, the second smaller
        print(f"noalloc case op={op} layout={layout} {s} cfg={cfg} splits={splits}", flush=True)
        in_, x, out0, epi, ref = _operands(op, layout, s, rng)
        cg.force_config(cfg, splits)
        try:
            got = _run(cg.conv_noalloc, op, layout, s, (in_, x), out0, epi)
            plan_na = cg.last_plan()
            want = _run(cg.conv_ex, op, layout, s, (in_, x), out0, epi)
            plan_ex = cg.last_plan()
        finally:
            cg.force_config(-1, 0)
        assert plan_na[:2] == plan_ex[:2], (plan_na, plan_ex)
        assert np.array_equal(got, want), "allocation-free call differs from convgemm_conv_ex"
        err, _ = rel_err(got, ref)
        assert err <= TOL, err


def test_noalloc_default_plan_and_oversized_weights():
    rng = np.random.default_rng(5)
    s = Shape(2, 14, 14, 64, 64, 3, 3, 1, 1, 1, 1, 1, 1)
    print(f"noalloc default plan {s}", flush=True)
    in_, x, out0, epi, ref = _operands(cg.FWD, NHWC, s, rng)
    got = _run(cg.conv_noalloc, cg.FWD, NHWC, s, (in_, x), out0, epi)
    assert rel_err(got, ref)[0] <= TOL
    # weights whose repacked copy exceeds the scratch: a clear error, nothing launched
    big = Shape(1, 3, 3, 1024, 1024, 3, 3, 1, 1, 1, 1, 1, 1)
    assert big.c * big.kh * big.kw * big.kn * 4 > cg.noalloc_scratch_bytes()
    w = torch.zeros(1, device=DEV)
    y = torch.zeros(1, device=DEV)
    with pytest.raises(cg.ConvGemmError) as ei:
        cg.conv_noalloc(_desc(cg.FWD, NHWC, big), w, w, y)
    assert ei.value.status == 4 and "scratch" in str(ei.value)
