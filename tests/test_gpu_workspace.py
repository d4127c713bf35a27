"""GPU tests of the host layer's workspace and batch-chunking paths.

* Batch chunking: every kernel addresses its tensors through 32-bit buffer
  offsets, so the host splits a problem whose operands exceed 2 GiB along the
  batch (later weight-gradient chunks accumulate with beta = 1).  A test hook
  shrinks that window so small problems run as 2-3 chunks; all six
  operations, split-K 1 and 3 and stream-K, with alpha and beta != 0,
  against the FP64 oracle.  One real forward and weight gradient with a
  2.25 GB input checks the unshrunk window at sampled outputs of both chunks.
* The internal workspace (calls without a workspace) under graph capture: a
  captured call keeps working after the buffer has grown (outgrown buffers
  are retired, not freed), and a call that would have to grow it while the
  stream is being captured fails with CONVGEMM_ERR_WORKSPACE instead.
"""
import numpy as np
import pytest
import torch

import oracle
from oracle import NCHW, NHWC, Shape, sampled

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    import convgemm_amd as cg
    cg.lib()
    yield
    cg.force_config(-1, 0)
    cg.set_chunk_window(0)


from tests.gpu_util import DEV, dev, host, random_problem, rel_err  # noqa: E402
import convgemm_amd as cg  # noqa: E402

OPK = {(cg.FWD, NHWC): 0, (cg.FWD, NCHW): 1, (cg.WGRAD, NHWC): 2, (cg.WGRAD, NCHW): 3, (cg.DGRAD, NHWC): 4,
       (cg.DGRAD, NCHW): 5}


def _desc(op, layout, s):
    return cg.make_desc(op, layout, s.b, s.h, s.w, s.c, s.kn, s.kh, s.kw, s.vpad, s.hpad, s.vstr, s.hstr, s.vdil,
                        s.hdil)


def _cfg_64(op, layout):
    for i, t in enumerate(cg.tile_configs(cg.F32)):
        if t[:2] == (64, 64) and t[5] == 0 and cg.lib().cg_tilecfg_supports(cg.F32, i, OPK[(op, layout)], 1):
            return i
    raise AssertionError("no 64x64 general configuration")


def _img_bytes(op, s):
    """Largest per-image operand or output (FP32) of the operation: what the chunk size divides."""
    x, y = s.h * s.w * s.c * 4, s.ho * s.wo * s.kn * 4
    return max(x, y)


@pytest.mark.parametrize("splits", [1, 3, 0])
@pytest.mark.parametrize("layout", [NHWC, NCHW])
@pytest.mark.parametrize("op", [cg.FWD, cg.WGRAD, cg.DGRAD])
def test_batch_chunks_against_oracle(op, layout, splits):
    s = Shape(5, 9, 8, 16, 24, 3, 3, 1, 1, 2, 1, 1, 1)
    rng = np.random.default_rng(100 + 10 * op + 3 * layout + splits)
    X, W, dY, e = random_problem(rng, s, layout, epi=op == cg.FWD)
    window = 2 * _img_bytes(op, s) + 64  # two images per chunk: chunks of 2, 2, 1
    cfg = _cfg_64(op, layout)
    print(f"chunked op={op} layout={layout} splits={splits} cfg={cfg} window={window}", flush=True)
    alpha, beta = 0.75, 0.5
    cg.set_chunk_window(window)
    cg.force_config(cfg, splits)
    try:
        if op == cg.FWD:
            Y0 = rng.standard_normal(s.y_shape(layout)).astype(np.float32)
            out = dev(Y0)
            ed = {k: dev(v) for k, v in e.items() if k != "relu"}
            cg.conv_ex(_desc(op, layout, s), dev(W), dev(X), out, alpha=alpha, beta=beta, relu=True, **ed)
            ref = oracle.fwd(layout, s, W, X, alpha=alpha, beta=beta, Y0=Y0, **e)
        elif op == cg.WGRAD:
            W0 = rng.standard_normal(s.w_shape(layout)).astype(np.float32)
            out = dev(W0)
            cg.conv_ex(_desc(op, layout, s), dev(dY), dev(X), out, alpha=alpha, beta=beta)
            ref = oracle.wgrad(layout, s, dY, X, alpha=alpha, beta=beta, dW0=W0)
        else:
            X0 = rng.standard_normal(s.x_shape(layout)).astype(np.float32)
            out = dev(X0)
            cg.conv_ex(_desc(op, layout, s), dev(W), dev(dY), out, alpha=alpha, beta=beta)
            ref = oracle.dgrad(layout, s, W, dY, alpha=alpha, beta=beta, dX0=X0)
        got = host(out)
        nk = cg.last_plan()[2]
    finally:
        cg.force_config(-1, 0)
        cg.set_chunk_window(0)
    assert nk >= 3, nk  # several chunks were launched
    assert rel_err(got, ref)[0] <= TOL


def test_operand_over_2gib():
    """1x1 forward and weight gradient on a 2.25 GB input (700 x 56 x 56 x 256 FP32): two batch chunks."""
    s = Shape(700, 56, 56, 256, 64, 1, 1, 0, 0, 1, 1, 1, 1)
    assert s.b * s.h * s.w * s.c * 4 > 2 ** 31
    print("over-2GiB forward and weight gradient (b=700, 56x56x256 -> 64)", flush=True)
    g = torch.Generator(device="cuda").manual_seed(77)
    X = torch.randn(s.x_shape(NHWC), device="cuda", generator=g)
    W = torch.randn(s.w_shape(NHWC), device="cuda", generator=g) * (2.0 / s.K) ** 0.5
    Y = torch.full(s.y_shape(NHWC), float("nan"), device="cuda")
    cg.conv_ex(_desc(cg.FWD, NHWC, s), W, X, Y)
    torch.cuda.synchronize()
    chunk = (2 ** 31 - 1) // (s.h * s.w * s.c * 4)
    assert chunk < s.b
    rng = np.random.default_rng(3)
    bs = np.concatenate([rng.integers(0, chunk, 48), rng.integers(chunk, s.b, 48)])
    ys = np.stack([bs, rng.integers(0, s.ho, 96), rng.integers(0, s.wo, 96), rng.integers(0, s.kn, 96)], 1)
    # a 1x1 output depends on its own pixel only: the oracle on those pixels as a batch of 1x1 images
    ib = torch.as_tensor(ys[:, 0], device="cuda")
    ih, iw = torch.as_tensor(ys[:, 1], device="cuda"), torch.as_tensor(ys[:, 2], device="cuda")
    Xs = X[ib, ih, iw, :].reshape(96, 1, 1, s.c).cpu().numpy()
    s1 = Shape(96, 1, 1, s.c, s.kn, 1, 1)
    idx = np.stack([np.arange(96), np.zeros(96, int), np.zeros(96, int), ys[:, 3]], 1)
    ref = sampled.fwd_sampled(NHWC, s1, W.cpu().numpy(), Xs, idx)
    got = Y[ib, ih, iw, torch.as_tensor(ys[:, 3], device="cuda")].double().cpu().numpy()
    assert np.max(np.abs(got - ref)) <= TOL * np.max(np.abs(ref))
    del Y
    # weight gradient: the second chunk accumulates into the first chunk's dW (beta = 1)
    dY = torch.randn(s.y_shape(NHWC), device="cuda", generator=g)
    dW = torch.full(s.w_shape(NHWC), float("nan"), device="cuda")
    cg.conv_ex(_desc(cg.WGRAD, NHWC, s), dY, X, dW)
    torch.cuda.synchronize()
    assert cg.last_plan()[2] >= 2
    # a 1x1 weight gradient is the plain product X^T dY (the GEMM identity the oracle tests pin), here in FP64
    ref = (X.view(-1, s.c).double().t() @ dY.view(-1, s.kn).double()).cpu().numpy()
    got = dW.view(s.c, s.kn).double().cpu().numpy()
    assert np.max(np.abs(got - ref)) <= TOL * np.max(np.abs(ref))


def test_capture_without_workspace():
    s = Shape(2, 14, 14, 32, 48, 3, 3, 1, 1, 1, 1, 1, 1)
    rng = np.random.default_rng(8)
    X, W, _, _ = random_problem(rng, s, NHWC, epi=False)
    ref = oracle.fwd(NHWC, s, W, X)
    Xd, Wd = dev(X), dev(W)
    Y = torch.full(s.y_shape(NHWC), float("nan"), device=DEV)
    d = _desc(cg.FWD, NHWC, s)
    cfg = _cfg_64(cg.FWD, NHWC)
    cg.force_config(cfg, 3)  # repacked weights and split-K partials in the internal workspace
    try:
        cg.conv_ex(d, Wd, Xd, Y)  # eager: the internal buffer holds what this call needs
        torch.cuda.synchronize()
        print("capture a call without a workspace", flush=True)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            cg.conv_ex(d, Wd, Xd, Y, stream=torch.cuda.current_stream())
        # an eager call with far larger repacked weights (151 MB) grows the internal buffer
        big = Shape(1, 4, 4, 2048, 2048, 3, 3, 1, 1, 1, 1, 1, 1)
        Wb = torch.randn(big.w_shape(NHWC), device=DEV) * 0.01
        Xb = torch.randn(big.x_shape(NHWC), device=DEV)
        Yb = torch.empty(big.y_shape(NHWC), device=DEV)
        cg.conv_ex(_desc(cg.FWD, NHWC, big), Wb, Xb, Yb)
        torch.cuda.synchronize()
        for rep in range(3):  # the graph still owns the outgrown buffer: replays stay correct
            Y.fill_(float("nan"))
            torch.cuda.synchronize()
            graph.replay()
            torch.cuda.synchronize()
            assert rel_err(host(Y), ref)[0] <= TOL, rep
            cg.conv_ex(_desc(cg.FWD, NHWC, big), Wb, Xb, Yb)  # eager work between replays
            torch.cuda.synchronize()
        # under capture a call that needs a larger buffer fails instead of growing it
        huge = Shape(1, 3, 3, 4096, 2048, 3, 3, 1, 1, 1, 1, 1, 1)  # 302 MB of repacked weights
        status = None
        g2 = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g2):
            try:
                cg.conv_ex(_desc(cg.FWD, NHWC, huge), Wb, Xb, Yb, stream=torch.cuda.current_stream())
            except cg.ConvGemmError as e:
                status = e.status
            Yb.mul_(1.0)  # keep the captured graph non-empty
        assert status == 4, status
    finally:
        cg.force_config(-1, 0)
