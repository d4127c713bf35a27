"""convgemm_amd -- MI355X (gfx950) convolution as implicit GEMM.

Python mirror of convGemm's C interface (src/convGemm.h:20-64) over the
C-ABI library ``libconvgemm.so`` (include/convgemm.h).  The functions take
torch tensors on the GPU (or raw device addresses as ints) and call the HIP
kernels through ctypes; torch is only used for device memory and streams.

The drop-in functions keep the reference's names, argument order and
meaning.  Where the reference has no error reporting (it returns void and
does no validation), these wrappers raise :class:`ConvGemmError` when the
library records a failure.

There is no CPU fallback: if ``libconvgemm.so`` is missing this module
raises on first use.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

__all__ = [
    "ConvGemmError", "lib", "alloc_pack_buffs", "free_pack_buffs", "sconvGemmNHWC", "sconvGemmNHWC_back",
    "sconvGemmNCHW", "sconvGemmNCHW_back", "conv_ex", "workspace_size", "out_size", "force_config",
    "FWD", "WGRAD", "DGRAD", "NHWC", "NCHW", "F32", "BF16", "FP8_E4M3", "EXPORTED_SYMBOLS",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
# CONVGEMM_LIB names another build of the same library (the sanitizer test runs the
# host-layer checks against an ASan/UBSan build this way)
LIB_PATH = os.environ.get("CONVGEMM_LIB") or os.path.join(_HERE, "libconvgemm.so")

FWD, WGRAD, DGRAD = 0, 1, 2
NHWC, NCHW = 0, 1
F32, BF16, FP8_E4M3 = 0, 1, 2

STATUS = {0: "success", 1: "invalid value", 2: "device allocation failed", 3: "HIP runtime error",
          4: "workspace too small", 5: "unsupported problem"}

# Every symbol include/convgemm.h declares.
EXPORTED_SYMBOLS = (
    "alloc_pack_buffs", "sconvGemmNHWC", "sconvGemmNHWC_back", "sconvGemmNCHW", "sconvGemmNCHW_back",
    "convgemm_conv_ex", "convgemm_workspace_size", "convgemm_out_size", "convgemm_free_pack_buffs",
    "convgemm_set_stream", "convgemm_get_last_error", "convgemm_last_error_message", "convgemm_status_string",
    "convgemm_version", "convgemm_quantize", "convgemm_packed_weights_size", "convgemm_pack_weights",
    "convgemm_conv_packed", "convgemm_conv_noalloc", "convgemm_noalloc_scratch_bytes",
    # convgemm_tune.h
    "convgemm_set_planner", "convgemm_get_planner", "convgemm_autotune", "convgemm_tune_count", "convgemm_tune_save",
    "convgemm_tune_load",
    "convgemm_tune_clear", "convgemm_plan", "convgemm_costmodel_train", "convgemm_costmodel_set_progress",
    "convgemm_costmodel_save",
    "convgemm_costmodel_load", "convgemm_costmodel_predict",
    # convgemm_batch.h
    "convgemm_batch_create", "convgemm_batch_run", "convgemm_batch_info", "convgemm_batch_plan",
    "convgemm_batch_destroy",
    # convgemm_explicit.h
    "im2row_nhwc", "row2im_nhwc", "im2col_nchw", "col2im_nchw", "transpose_nchw", "convgemm_gemm",
    "convgemm_im2row_ex", "convgemm_transpose_nchw_ex",
    "convgemm_nchw_to_nhwc", "convgemm_nhwc_to_nchw", "convgemm_weights_nhwc_to_nchw",
    "convgemm_weights_nchw_to_nhwc",
)

PLAN_ANALYTIC, PLAN_TUNED, PLAN_LEARNED, PLAN_MODEL = 0, 1, 2, 3
# CONVGEMM_MAX_CANDIDATES of convgemm_tune.h: candidate plans of one problem at most
MAX_CANDIDATES = 4096


class ConvGemmError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"convgemm: {STATUS.get(status, status)}: {message}")
        self.status = status


class Desc(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("op", "layout", "b", "h", "w", "c", "kn", "kh", "kw", "vpad", "hpad",
                                             "vstride", "hstride", "vdil", "hdil")]


class Epilogue(ctypes.Structure):
    _fields_ = [("bias", ctypes.c_void_p), ("mean", ctypes.c_void_p), ("inv_std", ctypes.c_void_p),
                ("gamma", ctypes.c_void_p), ("beta", ctypes.c_void_p), ("relu", ctypes.c_int)]


class Problem(ctypes.Structure):
    """convgemm_problem_t: one convolution of a batch."""
    _fields_ = [("desc", Desc), ("alpha", ctypes.c_float), ("beta", ctypes.c_float), ("in_", ctypes.c_void_p),
                ("x", ctypes.c_void_p), ("out", ctypes.c_void_p), ("epi", Epilogue),
                ("scales", ctypes.c_float * 2)]


BATCH_GRAPH, BATCH_PERSISTENT, BATCH_STATIC_WEIGHTS, BATCH_BOUNDED, BATCH_SINGLE = 1, 2, 4, 8, 16

_lib = None


def lib() -> ctypes.CDLL:
    """Load libconvgemm.so (after torch, so both share one HIP runtime)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m convgemm_amd.build` "
                          "(there is no CPU fallback)")
    try:  # torch ships libamdhip64.so.7; loading it first makes ours bind to the same runtime
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, ci, cf = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
    L.alloc_pack_buffs.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(vp)]
    L.alloc_pack_buffs.restype = ci
    L.sconvGemmNHWC.argtypes = [ctypes.c_char] + [ci] * 13 + [vp] * 8 + [ctypes.c_bool, vp, vp]
    L.sconvGemmNHWC.restype = None
    L.sconvGemmNCHW.argtypes = L.sconvGemmNHWC.argtypes
    L.sconvGemmNCHW.restype = None
    L.sconvGemmNHWC_back.argtypes = [ci] * 13 + [vp] * 5
    L.sconvGemmNHWC_back.restype = None
    L.sconvGemmNCHW_back.argtypes = L.sconvGemmNHWC_back.argtypes
    L.sconvGemmNCHW_back.restype = None
    L.convgemm_conv_ex.argtypes = [ctypes.POINTER(Desc), cf, cf, vp, vp, vp, ctypes.POINTER(Epilogue), ci, vp, vp,
                                   ctypes.c_size_t, vp]
    L.convgemm_conv_ex.restype = ci
    L.convgemm_packed_weights_size.argtypes = [ctypes.POINTER(Desc), ci]
    L.convgemm_packed_weights_size.restype = ctypes.c_size_t
    L.convgemm_pack_weights.argtypes = [ctypes.POINTER(Desc), ci, vp, vp, vp]
    L.convgemm_pack_weights.restype = ci
    L.convgemm_conv_packed.argtypes = L.convgemm_conv_ex.argtypes
    L.convgemm_conv_packed.restype = ci
    L.convgemm_workspace_size.argtypes = [ctypes.POINTER(Desc), ci]
    L.convgemm_workspace_size.restype = ctypes.c_size_t
    L.convgemm_out_size.argtypes = [ci] * 5
    L.convgemm_out_size.restype = ci
    L.convgemm_free_pack_buffs.argtypes = [vp, vp]
    L.convgemm_free_pack_buffs.restype = None
    L.convgemm_set_stream.argtypes = [vp]
    L.convgemm_set_stream.restype = ci
    L.convgemm_get_last_error.argtypes = []
    L.convgemm_get_last_error.restype = ci
    L.convgemm_last_error_message.argtypes = []
    L.convgemm_last_error_message.restype = ctypes.c_char_p
    L.convgemm_status_string.argtypes = [ci]
    L.convgemm_status_string.restype = ctypes.c_char_p
    L.convgemm_version.argtypes = []
    L.convgemm_version.restype = ci
    L.convgemm_quantize.argtypes = [vp, vp, ctypes.c_size_t, ci, ctypes.POINTER(cf), vp]
    L.convgemm_quantize.restype = ci
    L.convgemm_set_planner.argtypes = [ci]
    L.convgemm_set_planner.restype = None
    L.convgemm_get_planner.argtypes = []
    L.convgemm_get_planner.restype = ci
    L.convgemm_autotune.argtypes = [ctypes.POINTER(Desc), ci, vp, vp, vp, vp, ci, ctypes.POINTER(ci),
                                    ctypes.POINTER(ci), ctypes.POINTER(cf), vp]
    L.convgemm_autotune.restype = ci
    L.convgemm_tune_count.argtypes = []
    L.convgemm_tune_count.restype = ci
    L.convgemm_tune_save.argtypes = [ctypes.c_char_p]
    L.convgemm_tune_save.restype = ci
    L.convgemm_tune_load.argtypes = [ctypes.c_char_p]
    L.convgemm_tune_load.restype = ci
    L.convgemm_tune_clear.argtypes = []
    L.convgemm_tune_clear.restype = None
    L.convgemm_plan.argtypes = [ctypes.POINTER(Desc), ci, ctypes.POINTER(ci), ctypes.POINTER(ci)]
    L.convgemm_plan.restype = ci
    L.convgemm_costmodel_train.argtypes = [ci, cf, ctypes.c_uint, ctypes.POINTER(cf), vp]
    L.convgemm_costmodel_train.restype = ci
    L.convgemm_costmodel_set_progress.argtypes = [_TRAIN_PROGRESS, ci, vp]
    L.convgemm_costmodel_set_progress.restype = None
    L.convgemm_costmodel_save.argtypes = [ctypes.c_char_p]
    L.convgemm_costmodel_save.restype = ci
    L.convgemm_costmodel_load.argtypes = [ctypes.c_char_p]
    L.convgemm_costmodel_load.restype = ci
    L.convgemm_costmodel_predict.argtypes = [ctypes.POINTER(Desc), ci, ctypes.POINTER(ci), ctypes.POINTER(ci),
                                             ctypes.POINTER(cf), ctypes.POINTER(ci)]
    L.convgemm_costmodel_predict.restype = ci
    L.cg_tune_get.argtypes = [ctypes.POINTER(Desc), ci, ctypes.POINTER(ci), ctypes.POINTER(ci), ctypes.POINTER(cf), ci]
    L.cg_tune_get.restype = ci
    L.cg_mlp_loss_grad.argtypes = [ctypes.POINTER(cf), ctypes.POINTER(cf), ctypes.POINTER(cf), ci,
                                   ctypes.POINTER(cf), ctypes.POINTER(cf)]
    L.cg_mlp_loss_grad.restype = ci
    L.cg_mlp_dims.argtypes = [ctypes.POINTER(ci)] * 3
    L.cg_mlp_dims.restype = ci
    L.convgemm_batch_create.argtypes = [ctypes.POINTER(Problem), ci, ci, ctypes.c_uint, ctypes.POINTER(vp)]
    L.convgemm_batch_create.restype = ci
    L.convgemm_batch_run.argtypes = [vp, vp]
    L.convgemm_batch_run.restype = ci
    L.convgemm_batch_info.argtypes = [vp, ctypes.POINTER(ci), ctypes.POINTER(ci), ctypes.POINTER(ci)]
    L.convgemm_batch_info.restype = ci
    L.convgemm_batch_plan.argtypes = [vp, ci, ctypes.POINTER(ci), ctypes.POINTER(ci)]
    L.convgemm_batch_plan.restype = ci
    L.convgemm_batch_destroy.argtypes = [vp]
    L.convgemm_batch_destroy.restype = None
    L.im2row_nhwc.argtypes = [vp, ci, vp] + [ci] * 14
    L.im2row_nhwc.restype = None
    L.im2col_nchw.argtypes = L.im2row_nhwc.argtypes
    L.im2col_nchw.restype = None
    L.row2im_nhwc.argtypes = [ci, ci, vp, ci, vp] + [ci] * 14
    L.row2im_nhwc.restype = None
    L.col2im_nchw.argtypes = L.row2im_nhwc.argtypes
    L.col2im_nchw.restype = None
    L.transpose_nchw.argtypes = [ci, ci, vp, ci, cf, vp, ci, ci, ci, ci, ci]
    L.transpose_nchw.restype = None
    ch = ctypes.c_char
    L.convgemm_gemm.argtypes = [ch, ch, ch, ch, ch, ci, ci, ci, cf, vp, ci, vp, ci, cf, vp, ci, ci, vp, vp]
    L.convgemm_gemm.restype = ci
    for f in ("convgemm_nchw_to_nhwc", "convgemm_nhwc_to_nchw", "convgemm_weights_nhwc_to_nchw",
              "convgemm_weights_nchw_to_nhwc"):
        getattr(L, f).argtypes = [vp, vp, ci, ci, ci, ci, ci, vp]
        getattr(L, f).restype = ci
    L.cg_dgrad_class_order.argtypes = [ctypes.POINTER(Desc), ctypes.POINTER(ci), ctypes.POINTER(ci)]
    L.cg_dgrad_class_order.restype = ci
    L.cg_split_factors.argtypes = [ctypes.POINTER(ci), ci]
    L.cg_split_factors.restype = ci
    L.cg_num_tilecfgs.argtypes = [ci]
    L.cg_num_tilecfgs.restype = ci
    L.cg_get_tilecfg.argtypes = [ci, ci, ctypes.POINTER(ci * 8)]
    L.cg_num_tile_shapes.argtypes = [ci]
    L.cg_num_tile_shapes.restype = ci
    L.cg_get_tilecfg.restype = ci
    L.cg_tilecfg_supports.argtypes = [ci, ci, ci, ci]
    L.cg_tilecfg_supports.restype = ci
    L.cg_force_config.argtypes = [ci, ci]
    L.cg_force_config.restype = None
    L.cg_last_plan.argtypes = [ctypes.POINTER(ci)] * 3
    L.cg_last_plan.restype = None
    L.cg_debug_fail_next_allocs.argtypes = [ci]
    L.cg_debug_fail_next_allocs.restype = None
    L.convgemm_im2row_ex.argtypes = [ctypes.POINTER(Desc), ci, vp, vp, ctypes.c_longlong, vp]
    L.convgemm_im2row_ex.restype = ci
    L.convgemm_transpose_nchw_ex.argtypes = [ci, ci, vp, ctypes.c_longlong, cf, vp, ci, ci, ci, ci, ci, vp]
    L.convgemm_transpose_nchw_ex.restype = ci
    L.cg_set_chunk_window.argtypes = [ctypes.c_longlong]
    L.cg_set_chunk_window.restype = None
    L.cg_set_dgrad_forward.argtypes = [ci]
    L.cg_set_dgrad_forward.restype = None
    L.convgemm_conv_noalloc.argtypes = [ctypes.POINTER(Desc), cf, cf, vp, vp, vp, ctypes.POINTER(Epilogue), ci, vp,
                                        vp]
    L.convgemm_conv_noalloc.restype = ci
    L.convgemm_noalloc_scratch_bytes.argtypes = []
    L.convgemm_noalloc_scratch_bytes.restype = ctypes.c_size_t
    _lib = L
    return L


def _ptr(t) -> Optional[int]:
    if t is None:
        return None
    if isinstance(t, int):
        return t
    if hasattr(t, "data_ptr"):
        if not t.is_cuda:
            raise ValueError("convgemm_amd operates on GPU tensors only")
        if t.numel() == 0:
            return None
        if not t.is_contiguous():
            raise ValueError("tensors must be contiguous")
        return t.data_ptr()
    raise TypeError(f"expected a torch tensor or a device address, got {type(t)}")


def _check(L=None):
    L = L or lib()
    st = L.convgemm_get_last_error()
    if st != 0:
        raise ConvGemmError(st, L.convgemm_last_error_message().decode())


def last_error():
    L = lib()
    return L.convgemm_get_last_error(), L.convgemm_last_error_message().decode()


def out_size(inp, pad, dil, k, stride) -> int:
    """(in + 2*pad - dil*(k-1) - 1)/stride + 1 with C division (src/convGemm.c:60-61)."""
    return lib().convgemm_out_size(inp, pad, dil, k, stride)


def alloc_pack_buffs():
    """Mirror of alloc_pack_buffs (src/convGemm.c:30-40): returns (Ac_pack, Bc_pack) device addresses."""
    L = lib()
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if L.alloc_pack_buffs(ctypes.byref(a), ctypes.byref(b)) != 0:
        raise ConvGemmError(2, L.convgemm_last_error_message().decode())
    return a.value, b.value


def free_pack_buffs(ac, bc) -> None:
    lib().convgemm_free_pack_buffs(ac, bc)


def sconvGemmNHWC(trans, b, h, w, c, kn, kh, kw, vpadding, hpadding, vstride, hstride, vdilation, hdilation,
                  in_, x, out, bias_vector=None, bn_running_mean=None, bn_inv_std=None, bn_gamma=None,
                  bn_beta=None, relu=False, ac_pack=None, bc_pack=None):
    """trans=='N': out = epi(conv(x, in_)); else out = dW = im2row(x)^T in_ (src/convGemm.c:48-77)."""
    L = lib()
    L.sconvGemmNHWC(trans.encode() if isinstance(trans, str) else trans, b, h, w, c, kn, kh, kw, vpadding,
                    hpadding, vstride, hstride, vdilation, hdilation, _ptr(in_), _ptr(x), _ptr(out),
                    _ptr(bias_vector), _ptr(bn_running_mean), _ptr(bn_inv_std), _ptr(bn_gamma), _ptr(bn_beta),
                    bool(relu), _ptr(ac_pack), _ptr(bc_pack))
    _check(L)


def sconvGemmNHWC_back(b, h, w, c, kn, kh, kw, vstride, hstride, vpadding, hpadding, vdilation, hdilation,
                       weights, dy, dx, ac_pack=None, bc_pack=None):
    """dx += row2im(weights^T dy) (src/convGemm.c:82-97).  Stride before padding, as in the reference."""
    L = lib()
    L.sconvGemmNHWC_back(b, h, w, c, kn, kh, kw, vstride, hstride, vpadding, hpadding, vdilation, hdilation,
                         _ptr(weights), _ptr(dy), _ptr(dx), _ptr(ac_pack), _ptr(bc_pack))
    _check(L)


def sconvGemmNCHW(trans, b, c, h, w, kn, kh, kw, vpadding, hpadding, vstride, hstride, vdilation, hdilation,
                  in_, x, out, bias_vector=None, bn_running_mean=None, bn_inv_std=None, bn_gamma=None,
                  bn_beta=None, relu=False, ac_pack=None, bc_pack=None):
    """NCHW forward / weight gradient (src/convGemm.c:99-125)."""
    L = lib()
    L.sconvGemmNCHW(trans.encode() if isinstance(trans, str) else trans, b, c, h, w, kn, kh, kw, vpadding,
                    hpadding, vstride, hstride, vdilation, hdilation, _ptr(in_), _ptr(x), _ptr(out),
                    _ptr(bias_vector), _ptr(bn_running_mean), _ptr(bn_inv_std), _ptr(bn_gamma), _ptr(bn_beta),
                    bool(relu), _ptr(ac_pack), _ptr(bc_pack))
    _check(L)


def sconvGemmNCHW_back(b, c, h, w, kn, kh, kw, vstride, hstride, vpadding, hpadding, vdilation, hdilation,
                       weights, dy, dx, ac_pack=None, bc_pack=None):
    """dx += col2im(dy weights) (src/convGemm.c:130-146)."""
    L = lib()
    L.sconvGemmNCHW_back(b, c, h, w, kn, kh, kw, vstride, hstride, vpadding, hpadding, vdilation, hdilation,
                         _ptr(weights), _ptr(dy), _ptr(dx), _ptr(ac_pack), _ptr(bc_pack))
    _check(L)


def make_desc(op, layout, b, h, w, c, kn, kh, kw, vpad=0, hpad=0, vstride=1, hstride=1, vdil=1, hdil=1) -> Desc:
    return Desc(op, layout, b, h, w, c, kn, kh, kw, vpad, hpad, vstride, hstride, vdil, hdil)


def conv_ex(desc: Desc, in_, x, out, alpha=1.0, beta=0.0, bias=None, mean=None, inv_std=None, gamma=None,
            bnbeta=None, relu=False, dtype=F32, scales=None, workspace=None, ws_bytes=0, stream=None):
    """convgemm_conv_ex: asynchronous on `stream` (torch stream, raw handle or None = null stream)."""
    L = lib()
    epi = Epilogue(_ptr(bias), _ptr(mean), _ptr(inv_std), _ptr(gamma), _ptr(bnbeta), int(bool(relu)))
    if stream is None:
        sh = None
    elif isinstance(stream, int):
        sh = stream
    else:
        sh = stream.cuda_stream
    sc = None
    if scales is not None:  # host floats {scale_in, scale_x}
        sc = ctypes.cast((ctypes.c_float * 2)(*[float(v) for v in scales]), ctypes.c_void_p)
    st = L.convgemm_conv_ex(ctypes.byref(desc), float(alpha), float(beta), _ptr(in_), _ptr(x), _ptr(out),
                            ctypes.byref(epi), int(dtype), sc, _ptr(workspace), int(ws_bytes), sh)
    if st != 0:
        raise ConvGemmError(st, L.convgemm_last_error_message().decode())


def conv_noalloc(desc: Desc, in_, x, out, alpha=1.0, beta=0.0, bias=None, mean=None, inv_std=None, gamma=None,
                 bnbeta=None, relu=False, dtype=F32, scales=None, stream=None):
    """convgemm_conv_noalloc: conv_ex on the library's static device scratch (no device allocation);
    calls that overlap in time on one GPU share that scratch and race."""
    L = lib()
    epi = Epilogue(_ptr(bias), _ptr(mean), _ptr(inv_std), _ptr(gamma), _ptr(bnbeta), int(bool(relu)))
    st = L.convgemm_conv_noalloc(ctypes.byref(desc), float(alpha), float(beta), _ptr(in_), _ptr(x), _ptr(out),
                                 ctypes.byref(epi), int(dtype), _scales_ptr(scales), _stream_handle(stream))
    if st != 0:
        raise ConvGemmError(st, L.convgemm_last_error_message().decode())


def noalloc_scratch_bytes() -> int:
    """Size of the static device scratch behind conv_noalloc."""
    return int(lib().convgemm_noalloc_scratch_bytes())


def im2row_ex(desc: Desc, x, out, ld, dtype=F32, stream=None) -> None:
    """convgemm_im2row_ex: the patch matrix of desc's input (NHWC rows [P][ld] / NCHW cols [K][ld]) in
    FP32 or BF16 elements, asynchronous on `stream`; only in-bounds taps are written."""
    L = lib()
    st = L.convgemm_im2row_ex(ctypes.byref(desc), int(dtype), _ptr(x), _ptr(out), int(ld), _stream_handle(stream))
    if st != 0:
        raise ConvGemmError(st, L.convgemm_last_error_message().decode())


def transpose_nchw_ex(rows, cols, in_, ld, beta, out, kn, ho, wo, start_row=0, start_col=0, stream=None) -> None:
    """convgemm_transpose_nchw_ex: transpose_nchw asynchronous on `stream`."""
    L = lib()
    st = L.convgemm_transpose_nchw_ex(rows, cols, _ptr(in_), int(ld), float(beta), _ptr(out), kn, ho, wo, start_row,
                                      start_col, _stream_handle(stream))
    if st != 0:
        raise ConvGemmError(st, L.convgemm_last_error_message().decode())


def set_chunk_window(nbytes: int) -> None:
    """Test hook: bytes one batch chunk may address (0 restores 2 GiB); smaller values make small
    problems run the batch-chunking path."""
    lib().cg_set_chunk_window(int(nbytes))


def set_dgrad_forward(on: bool) -> None:
    """Test hook: backward data of unit-stride NHWC problems runs as the forward convolution of dY
    with the kernel turned by 180 degrees (True, the default) or on the residue-class kernel."""
    lib().cg_set_dgrad_forward(1 if on else 0)


def quantize(src, dst, dtype, stream=None) -> float:
    """FP32 -> BF16 (returns 1.0) or FP8 e4m3 with a per-tensor amax scale (returns the scale)."""
    L = lib()
    sc = ctypes.c_float(1.0)
    sh = None if stream is None else (stream if isinstance(stream, int) else stream.cuda_stream)
    st = L.convgemm_quantize(_ptr(src), _ptr(dst), int(src.numel()), int(dtype), ctypes.byref(sc), sh)
    if st != 0:
        raise ConvGemmError(st, L.convgemm_last_error_message().decode())
    return sc.value


def workspace_size(desc: Desc, dtype=F32) -> int:
    return int(lib().convgemm_workspace_size(ctypes.byref(desc), int(dtype)))


def _stream_handle(stream):
    if stream is None:
        return None
    return stream if isinstance(stream, int) else stream.cuda_stream


def _scales_ptr(scales):
    if scales is None:
        return None
    return ctypes.cast((ctypes.c_float * 2)(*[float(v) for v in scales]), ctypes.c_void_p)


def set_planner(mode: int) -> None:
    """PLAN_ANALYTIC, PLAN_TUNED, PLAN_LEARNED or PLAN_MODEL.  The library starts with
    PLAN_LEARNED when it loaded a cost model at start-up (the shipped one, or the file
    CONVGEMM_COSTMODEL names), PLAN_TUNED otherwise."""
    lib().convgemm_set_planner(int(mode))


def get_planner() -> int:
    return lib().convgemm_get_planner()


def autotune(desc: Desc, in_, x, out, dtype=F32, scales=None, reps=3, stream=None):
    """Time every (tile configuration, split-K) candidate on these buffers; returns (cfg, splits, us)."""
    L = lib()
    c, s, us = ctypes.c_int(), ctypes.c_int(), ctypes.c_float()
    st = L.convgemm_autotune(ctypes.byref(desc), int(dtype), _scales_ptr(scales), _ptr(in_), _ptr(x), _ptr(out),
                             int(reps), ctypes.byref(c), ctypes.byref(s), ctypes.byref(us), _stream_handle(stream))
    if st != 0:
        raise ConvGemmError(st, L.convgemm_last_error_message().decode())
    return c.value, s.value, us.value


def plan(desc: Desc, dtype=F32):
    """(tile configuration, split-K) the library would use for this problem now."""
    L = lib()
    c, s = ctypes.c_int(), ctypes.c_int()
    st = L.convgemm_plan(ctypes.byref(desc), int(dtype), ctypes.byref(c), ctypes.byref(s))
    if st != 0:
        raise ConvGemmError(st, L.convgemm_last_error_message().decode())
    return c.value, s.value


def tune_count() -> int:
    return int(lib().convgemm_tune_count())


def tune_save(path: str) -> None:
    if lib().convgemm_tune_save(path.encode()) != 0:
        raise ConvGemmError(1, lib().convgemm_last_error_message().decode())


def tune_load(path: str) -> None:
    if lib().convgemm_tune_load(path.encode()) != 0:
        raise ConvGemmError(1, lib().convgemm_last_error_message().decode())


def tune_clear() -> None:
    lib().convgemm_tune_clear()


def tune_get(desc: Desc, dtype=F32):
    """Recorded timings of a tuned problem: list of (cfg, splits, microseconds)."""
    L = lib()
    m = MAX_CANDIDATES
    c, s, u = (ctypes.c_int * m)(), (ctypes.c_int * m)(), (ctypes.c_float * m)()
    n = L.cg_tune_get(ctypes.byref(desc), int(dtype), c, s, u, m)
    return [(c[i], s[i], u[i]) for i in range(n)]


_TRAIN_PROGRESS = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_float, ctypes.c_void_p)


def costmodel_train(epochs=1500, lr=3e-3, seed=0, stream=None, progress=None, every=100) -> float:
    """Train the cost model on every timing in the tuning database (device kernels); returns the final loss.
    progress(epoch, loss), if given, is called every `every` epochs."""
    L = lib()
    loss = ctypes.c_float()
    cb = _TRAIN_PROGRESS(lambda ep, l, _ctx: progress(ep, l)) if progress else _TRAIN_PROGRESS()
    L.convgemm_costmodel_set_progress(cb, int(every) if progress else 0, None)
    try:
        st = L.convgemm_costmodel_train(int(epochs), float(lr), int(seed), ctypes.byref(loss),
                                        _stream_handle(stream))
    finally:
        L.convgemm_costmodel_set_progress(_TRAIN_PROGRESS(), 0, None)
    if st != 0:
        raise ConvGemmError(st, L.convgemm_last_error_message().decode())
    return loss.value


def costmodel_save(path: str) -> None:
    if lib().convgemm_costmodel_save(path.encode()) != 0:
        raise ConvGemmError(1, lib().convgemm_last_error_message().decode())


def costmodel_load(path: str) -> None:
    if lib().convgemm_costmodel_load(path.encode()) != 0:
        raise ConvGemmError(1, lib().convgemm_last_error_message().decode())


def costmodel_predict(desc: Desc, dtype=F32):
    """Predicted microseconds of every candidate: list of (cfg, splits, us)."""
    L = lib()
    m = MAX_CANDIDATES
    c, s, u = (ctypes.c_int * m)(), (ctypes.c_int * m)(), (ctypes.c_float * m)()
    n = ctypes.c_int(m)
    st = L.convgemm_costmodel_predict(ctypes.byref(desc), int(dtype), c, s, u, ctypes.byref(n))
    if st != 0:
        raise ConvGemmError(st, L.convgemm_last_error_message().decode())
    return [(c[i], s[i], u[i]) for i in range(n.value)]


def mlp_loss_grad(params, X, T):
    """Loss and gradients of the cost-model MLP computed by the device kernels (test hook)."""
    import numpy as np
    L = lib()
    F, H, NP = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    L.cg_mlp_dims(ctypes.byref(F), ctypes.byref(H), ctypes.byref(NP))
    p = np.ascontiguousarray(params, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    T = np.ascontiguousarray(T, np.float32)
    assert p.size == NP.value and X.shape[1] == F.value and X.shape[0] == T.size
    g = np.empty(NP.value, np.float32)
    loss = ctypes.c_float()
    fp = ctypes.POINTER(ctypes.c_float)
    rc = L.cg_mlp_loss_grad(p.ctypes.data_as(fp), X.ctypes.data_as(fp), T.ctypes.data_as(fp), int(T.size),
                            g.ctypes.data_as(fp), ctypes.byref(loss))
    if rc != 0:
        raise ConvGemmError(3, "mlp kernels failed")
    return float(loss.value), g


def last_plan():
    """(tile configuration, split-K factor, kernels launched) of the last call on this thread."""
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib().cg_last_plan(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    return a.value, b.value, c.value


def force_config(cfg: int = -1, splits: int = 0) -> None:
    """Tuning hook: force tile configuration `cfg` and split-K `splits` on every call (cfg < 0 clears)."""
    lib().cg_force_config(int(cfg), int(splits))


def problem(desc: Desc, in_, x, out, alpha=1.0, beta=0.0, bias=None, mean=None, inv_std=None, gamma=None,
            bnbeta=None, relu=False, scales=None) -> Problem:
    """One convolution for a Batch (the arguments of conv_ex)."""
    pr = Problem()
    pr.desc = desc
    pr.alpha, pr.beta = float(alpha), float(beta)
    pr.in_, pr.x, pr.out = _ptr(in_), _ptr(x), _ptr(out)
    pr.epi = Epilogue(_ptr(bias), _ptr(mean), _ptr(inv_std), _ptr(gamma), _ptr(bnbeta), int(bool(relu)))
    sc = scales if scales is not None else (1.0, 1.0)
    pr.scales[0], pr.scales[1] = float(sc[0]), float(sc[1])
    return pr


class Batch:
    """convgemm_batch_t: many independent convolutions in grouped launches,
    optionally replayed from a hipGraph (flags BATCH_GRAPH, BATCH_PERSISTENT,
    BATCH_STATIC_WEIGHTS, BATCH_BOUNDED, BATCH_SINGLE).  Keep the operand tensors alive while
    the batch is used: their addresses are bound at creation."""

    def __init__(self, problems, dtype=F32, flags=0, keep=None):
        L = lib()
        arr = (Problem * max(len(problems), 1))(*problems)
        h = ctypes.c_void_p()
        st = L.convgemm_batch_create(arr, len(problems), int(dtype), int(flags), ctypes.byref(h))
        if st != 0:
            raise ConvGemmError(st, L.convgemm_last_error_message().decode())
        self._h = h
        self._keep = keep

    def run(self, stream=None) -> None:
        L = lib()
        st = L.convgemm_batch_run(self._h, _stream_handle(stream))
        if st != 0:
            raise ConvGemmError(st, L.convgemm_last_error_message().decode())

    def info(self):
        """(grouped implicit-GEMM launches, kernels per run, replays a graph)."""
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        st = lib().convgemm_batch_info(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        if st != 0:
            raise ConvGemmError(st, lib().convgemm_last_error_message().decode())
        return a.value, b.value, bool(c.value)

    def plan(self, i):
        """(cfg, splits) of problem i as given at creation; cfg -1: no implicit GEMM."""
        a, b = ctypes.c_int(), ctypes.c_int()
        st = lib().convgemm_batch_plan(self._h, int(i), ctypes.byref(a), ctypes.byref(b))
        if st != 0:
            raise ConvGemmError(st, lib().convgemm_last_error_message().decode())
        return a.value, b.value

    def close(self) -> None:
        if self._h:
            lib().convgemm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- explicit path (include/convgemm_explicit.h) ----

def im2row_nhwc(rows, ld, in_, batches, height, width, channels, oheight, owidth, kheight, kwidth, vpadding,
                hpadding, vstride, hstride, vdilation, hdilation):
    """rows[p*ld + r] = in[b][ix][iy][c] for in-bounds taps (src/im2row_nhwc.c:147-204)."""
    L = lib()
    L.im2row_nhwc(_ptr(rows), ld, _ptr(in_), batches, height, width, channels, oheight, owidth, kheight, kwidth,
                  vpadding, hpadding, vstride, hstride, vdilation, hdilation)
    _check(L)


def row2im_nhwc(m, n, rows, ld, out, batches, height, width, channels, oheight, owidth, kheight, kwidth, vpadding,
                hpadding, vstride, hstride, vdilation, hdilation):
    """out += row2im(rows) (src/im2row_nhwc.c:206-229)."""
    L = lib()
    L.row2im_nhwc(m, n, _ptr(rows), ld, _ptr(out), batches, height, width, channels, oheight, owidth, kheight,
                  kwidth, vpadding, hpadding, vstride, hstride, vdilation, hdilation)
    _check(L)


def im2col_nchw(cols, ld, in_, batches, channels, height, width, oheight, owidth, kheight, kwidth, vpadding,
                hpadding, vstride, hstride, vdilation, hdilation):
    """cols[r*ld + p] = in[b][c][ix][iy] for in-bounds taps (src/im2col_nchw.c:146-204)."""
    L = lib()
    L.im2col_nchw(_ptr(cols), ld, _ptr(in_), batches, channels, height, width, oheight, owidth, kheight, kwidth,
                  vpadding, hpadding, vstride, hstride, vdilation, hdilation)
    _check(L)


def col2im_nchw(m, n, cols, ld, out, batches, channels, height, width, oheight, owidth, kheight, kwidth, vpadding,
                hpadding, vstride, hstride, vdilation, hdilation):
    """out += col2im(cols) (src/im2col_nchw.c:337-359)."""
    L = lib()
    L.col2im_nchw(m, n, _ptr(cols), ld, _ptr(out), batches, channels, height, width, oheight, owidth, kheight,
                  kwidth, vpadding, hpadding, vstride, hstride, vdilation, hdilation)
    _check(L)


def transpose_nchw(rows, cols, in_, ld, beta, out, kn, ho, wo, start_row, start_col):
    """[kn-block][b][ho][wo] -> [b][kn][ho][wo] with beta (src/im2col_nchw.c:206-233)."""
    L = lib()
    L.transpose_nchw(rows, cols, _ptr(in_), ld, float(beta), _ptr(out), kn, ho, wo, start_row, start_col)
    _check(L)


def gemm(orderA, orderB, orderC, transA, transB, m, n, k, alpha, A, ldA, B, ldB, beta, C, ldC, dtype=F32,
         scales=None, stream=None):
    """C = alpha*op(A)*op(B) + beta*C on the implicit-GEMM kernels (convgemm_gemm)."""
    L = lib()
    enc = lambda ch: ch.encode() if isinstance(ch, str) else ch  # noqa: E731
    st = L.convgemm_gemm(enc(orderA), enc(orderB), enc(orderC), enc(transA), enc(transB), m, n, k, float(alpha),
                         _ptr(A), ldA, _ptr(B), ldB, float(beta), _ptr(C), ldC, int(dtype), _scales_ptr(scales),
                         _stream_handle(stream))
    if st != 0:
        raise ConvGemmError(st, L.convgemm_last_error_message().decode())


def _conv_layout(fn, src, dst, a, b, c, d, esize, stream):
    L = lib()
    st = getattr(L, fn)(_ptr(src), _ptr(dst), a, b, c, d, esize, _stream_handle(stream))
    if st != 0:
        raise ConvGemmError(st, L.convgemm_last_error_message().decode())


def nchw_to_nhwc(src, dst, b, c, h, w, esize=4, stream=None):
    _conv_layout("convgemm_nchw_to_nhwc", src, dst, b, c, h, w, esize, stream)


def nhwc_to_nchw(src, dst, b, h, w, c, esize=4, stream=None):
    _conv_layout("convgemm_nhwc_to_nchw", src, dst, b, h, w, c, esize, stream)


def weights_nhwc_to_nchw(src, dst, c, kh, kw, kn, esize=4, stream=None):
    _conv_layout("convgemm_weights_nhwc_to_nchw", src, dst, c, kh, kw, kn, esize, stream)


def weights_nchw_to_nhwc(src, dst, kn, c, kh, kw, esize=4, stream=None):
    _conv_layout("convgemm_weights_nchw_to_nhwc", src, dst, kn, c, kh, kw, esize, stream)


def tile_configs(dtype=F32):
    """Configurations of an operand type: list of (BM, BN, BK, WM, WN, pipelined stages or 0,
    raster order, halo slots or 0).  Configuration base + r * num_tile_shapes(dtype) is tile
    shape `base` walked in raster order r (0: the N panels of an M panel adjacent, 1: the M
    panels of an N panel adjacent, 2: bands of 8 M panels).  A configuration with halo slots
    is the halo-tile forward kernel: its M tiles hold the pixels of one image, and it serves
    a problem only when the input rows those pixels read fit the slots."""
    L = lib()
    out = []
    for c in range(L.cg_num_tilecfgs(int(dtype))):
        v = (ctypes.c_int * 8)()
        if L.cg_get_tilecfg(int(dtype), c, ctypes.byref(v)) == 0:
            out.append(tuple(v))
    return out


def num_tile_shapes(dtype=F32) -> int:
    """Tile shapes of an operand type (configurations 0 .. n-1 are the shapes in raster order 0)."""
    return int(lib().cg_num_tile_shapes(int(dtype)))


def dgrad_class_order(desc: Desc):
    """Backward-data launch order of the stride residue classes (class = rh * hstride + rw) and
    each class's tap count; ([], []) when the order is the identity."""
    o, t = (ctypes.c_int * 16)(), (ctypes.c_int * 16)()
    n = lib().cg_dgrad_class_order(ctypes.byref(desc), o, t)
    return [o[i] for i in range(n)], [t[i] for i in range(n)]


def split_factors():
    """Split-K factors the tuner and the cost model try for every configuration."""
    buf = (ctypes.c_int * 64)()
    n = lib().cg_split_factors(buf, 64)
    return [buf[i] for i in range(min(n, 64))]


def tile_supports(dtype, cfg, op, layout, vec=True) -> bool:
    """Is tile configuration `cfg` instantiated for (op, layout), with 16-byte operand loads if `vec`?"""
    opkind = {(FWD, NHWC): 0, (FWD, NCHW): 1, (WGRAD, NHWC): 2, (WGRAD, NCHW): 3, (DGRAD, NHWC): 4,
              (DGRAD, NCHW): 5}[(op, layout)]
    return bool(lib().cg_tilecfg_supports(int(dtype), int(cfg), opkind, int(bool(vec))))


def packed_weights_size(desc: Desc, dtype=F32) -> int:
    """Bytes of the prepacked weights of this problem (0: weights are read as given)."""
    return int(lib().convgemm_packed_weights_size(ctypes.byref(desc), int(dtype)))


def pack_weights(desc: Desc, in_, packed, dtype=F32, stream=None) -> None:
    """Repack the weights once into `packed` for conv_packed."""
    L = lib()
    st = L.convgemm_pack_weights(ctypes.byref(desc), int(dtype), _ptr(in_), _ptr(packed), _stream_handle(stream))
    if st != 0:
        raise ConvGemmError(st, L.convgemm_last_error_message().decode())


def conv_packed(desc: Desc, packed, x, out, alpha=1.0, beta=0.0, bias=None, mean=None, inv_std=None, gamma=None,
                bnbeta=None, relu=False, dtype=F32, scales=None, workspace=None, ws_bytes=0, stream=None):
    """conv_ex on weights prepacked by pack_weights (no repack kernel per call)."""
    L = lib()
    epi = Epilogue(_ptr(bias), _ptr(mean), _ptr(inv_std), _ptr(gamma), _ptr(bnbeta), int(bool(relu)))
    st = L.convgemm_conv_packed(ctypes.byref(desc), float(alpha), float(beta), _ptr(packed), _ptr(x), _ptr(out),
                                ctypes.byref(epi), int(dtype), _scales_ptr(scales), _ptr(workspace), int(ws_bytes),
                                _stream_handle(stream))
    if st != 0:
        raise ConvGemmError(st, L.convgemm_last_error_message().decode())
