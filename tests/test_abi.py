"""C-ABI boundary checks that need no GPU.

* libconvgemm.so loads and exports every function include/*.h declares;
* the host layer rejects invalid problems before any device work, recording
  a status (the reference has no validation: src/convGemm.c:48-146);
* the fast-division helper used for pixel/tap decomposition in the kernels is
  exact (checked on the host, exhaustively over small divisors).
"""
import ctypes
import os
import re
import subprocess
import sys
import tempfile

import pytest

import convgemm_amd as cg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = set()
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", fn)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        txt = re.sub(r"typedef\s+\w+\s*\(\s*\*\s*\w+\s*\)\s*\([^;]*\);", "", txt)  # callback types
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt):
            name = m.group(1)
            if name in ("if", "sizeof", "defined"):
                continue
            # a declaration: preceded by a return type on the same statement
            names.add(name)
    return names


def test_header_symbols_match_exported_list():
    declared = header_functions()
    assert declared == set(cg.EXPORTED_SYMBOLS), declared ^ set(cg.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    L = cg.lib()
    for name in cg.EXPORTED_SYMBOLS:
        assert hasattr(L, name), name
    nm = subprocess.run(["nm", "-D", "--defined-only", cg.LIB_PATH], stdout=subprocess.PIPE, text=True).stdout
    defined = {l.split()[-1] for l in nm.splitlines() if " T " in l}
    missing = set(cg.EXPORTED_SYMBOLS) - defined
    assert not missing, missing


def test_drop_in_signatures_mirror_reference_order():
    """The NHWC forward takes padding before stride, _back takes stride before padding
    (src/convGemm.h:22-42); check the header text keeps that trap."""
    h = open(os.path.join(ROOT, "include", "convgemm.h")).read()
    fwd = h[h.index("void sconvGemmNHWC(char trans"):]
    fwd = fwd[:fwd.index(";")]
    assert fwd.index("vpadding") < fwd.index("vstride")
    back = h[h.index("void sconvGemmNHWC_back("):]
    back = back[:back.index(";")]
    assert back.index("vstride") < back.index("vpadding")


def test_version_and_status_strings():
    L = cg.lib()
    assert L.convgemm_version() >= 10000
    assert L.convgemm_status_string(0) == b"success"
    assert L.convgemm_status_string(4) == b"workspace too small"


def test_out_size_formula():
    assert cg.out_size(56, 1, 1, 3, 1) == 56
    assert cg.out_size(56, 1, 1, 3, 2) == 28
    assert cg.out_size(55, 0, 1, 11, 4) == 12
    assert cg.out_size(7, 2, 2, 3, 1) == 7


BAD = [
    dict(vstride=0),
    dict(hstride=-1),
    dict(vdil=0),
    dict(vpad=-1),
    dict(kh=0),
    dict(b=-1),
    dict(h=2, w=2, kh=5, kw=5),      # empty output
    dict(op=7),
    dict(layout=3),
]


@pytest.mark.parametrize("bad", BAD)
def test_conv_ex_rejects_invalid_descriptors(bad):
    d = dict(op=cg.FWD, layout=cg.NHWC, b=1, h=8, w=8, c=4, kn=4, kh=3, kw=3, vpad=1, hpad=1, vstride=1,
             hstride=1, vdil=1, hdil=1)
    d.update(bad)
    desc = cg.Desc(**d)
    with pytest.raises(cg.ConvGemmError) as e:
        cg.conv_ex(desc, 0x1000, 0x2000, 0x3000)
    assert e.value.status in (1, 5)


def test_drop_in_records_error_without_touching_device():
    L = cg.lib()
    # stride 0: rejected before any HIP call, nothing written, status recorded
    L.sconvGemmNHWC(b"N", 1, 8, 8, 4, 4, 3, 3, 1, 1, 0, 1, 1, 1, 0x1000, 0x2000, 0x3000,
                    None, None, None, None, None, False, None, None)
    st, msg = cg.last_error()
    assert st == 1 and "stride" in msg
    L.sconvGemmNCHW_back(1, 4, 8, 8, 4, 3, 3, 1, 1, 1, 1, 1, 1, None, None, 0x3000, None, None)
    st, msg = cg.last_error()
    assert st == 1 and "NULL" in msg


def test_batch_norm_requires_all_vectors():
    desc = cg.make_desc(cg.FWD, cg.NHWC, 1, 8, 8, 4, 4, 3, 3, 1, 1)
    with pytest.raises(cg.ConvGemmError) as e:
        cg.conv_ex(desc, 0x1000, 0x2000, 0x3000, mean=0x4000)  # inv_std/gamma/beta missing
    assert e.value.status == 1


def test_quick_return_needs_no_device():
    """m == 0 or n == 0 returns before any device work (B3A2C0_orig.c:45-47)."""
    for kw in (dict(b=0), dict(kn=0)):
        d = dict(op=cg.FWD, layout=cg.NHWC, b=1, h=8, w=8, c=4, kn=4, kh=3, kw=3, vpad=1, hpad=1, vstride=1,
                 hstride=1, vdil=1, hdil=1)
        d.update(kw)
        cg.conv_ex(cg.Desc(**d), None, None, None)
    # k == 0 (c == 0) with beta == 1: quick return
    d = cg.make_desc(cg.FWD, cg.NHWC, 1, 8, 8, 0, 4, 3, 3, 1, 1)
    cg.conv_ex(d, None, None, None, beta=1.0)


def test_alloc_failure_returns_one():
    L = cg.lib()
    L.cg_debug_fail_next_allocs(1)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert L.alloc_pack_buffs(ctypes.byref(a), ctypes.byref(b)) == 1
    assert a.value is None and b.value is None
    assert L.convgemm_get_last_error() == 2


FASTDIV_TEST = r"""
#include "kernels/igemm_common.hpp"
#include <cstdio>
#include <cstdlib>
int main() {
    unsigned long long checked = 0;
    for (int d = 1; d <= 4096; d++) {
        cg::FastDiv f = cg::make_fastdiv(d);
        for (int n = 0; n < 70000; n++) { if (cg::fdiv(n, f) != n / d) { printf("FAIL %d %d\n", n, d); return 1; } checked++; }
        int big[] = {2147483647, 2147483646, 2147483647 - d, 1 << 30, (1 << 30) + d * 7, 123456789};
        for (int n : big) { if (n < 0) continue; if (cg::fdiv(n, f) != n / d) { printf("FAIL %d %d\n", n, d); return 1; } }
    }
    int ds[] = {3136, 12544, 100352, 784, 49, 9, 25, 121, 2147483647, 1 << 30, 1000003};
    for (int d : ds) {
        cg::FastDiv f = cg::make_fastdiv(d);
        for (long long n = 0; n < 2147483648LL; n += 9973) { if (cg::fdiv((int)n, f) != (int)(n / d)) { printf("FAIL %lld %d\n", n, d); return 1; } }
    }
    printf("OK %llu\n", checked);
    return 0;
}
"""


def test_fastdiv_exact():
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "t.cpp")
        exe = os.path.join(td, "t")
        open(src, "w").write(FASTDIV_TEST)
        r = subprocess.run([hipcc, "-O2", "-x", "hip", "--offload-arch=gfx950", "-I",
                            os.path.join(ROOT, "convgemm_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
                            src, "-o", exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        assert r.returncode == 0, r.stdout
        out = subprocess.run([exe], stdout=subprocess.PIPE, text=True)
        assert out.returncode == 0 and out.stdout.startswith("OK"), out.stdout


def test_candidate_capacity_covers_every_tile_table():
    """Every configuration with every split-K factor, plus one stream-K plan per
    configuration, fits CONVGEMM_MAX_CANDIDATES: the autotuner and the cost
    model rank the whole candidate list of a problem, never a prefix of it."""
    hdr = open(os.path.join(ROOT, "include", "convgemm_tune.h")).read()
    m = re.search(r"#define\s+CONVGEMM_MAX_CANDIDATES\s+(\d+)", hdr)
    assert m and int(m.group(1)) == cg.MAX_CANDIDATES
    per_cfg = len(cg.split_factors()) + 1
    assert per_cfg > 2
    for dt in (cg.F32, cg.BF16, cg.FP8_E4M3):
        n = len(cg.tile_configs(dt))
        assert n > 0 and n * per_cfg <= cg.MAX_CANDIDATES, (dt, n, per_cfg)


def test_mix_spec_parsing():
    """--mix entries of the tuning module: dtype:op:layout with an optional training count."""
    from convgemm_amd.tuning import parse_mix
    m = parse_mix("f32:fwd:nchw,bf16:wgrad:nhwc:2000", 150)
    assert m == [("f32:fwd:nchw", cg.F32, cg.FWD, cg.NCHW, 150), ("bf16:wgrad:nhwc", cg.BF16, cg.WGRAD, cg.NHWC, 2000)]
    with pytest.raises(ValueError):
        parse_mix("bf16:fwd", 10)


SHIPPED_MODEL = os.path.join(ROOT, "convgemm_amd", "data", "mi355x_costmodel.mlp")


@pytest.mark.parametrize("env,expect", [(None, "PLAN_LEARNED"), ("none", "PLAN_TUNED"), ("", "PLAN_TUNED"),
                                        ("/nonexistent/model.mlp", "PLAN_TUNED"), (SHIPPED_MODEL, "PLAN_LEARNED")])
def test_startup_planner_follows_cost_model(env, expect):
    """The library loads the shipped cost model (or CONVGEMM_COSTMODEL's file) when it
    is loaded, and then plans with PLAN_LEARNED; without a model, PLAN_TUNED.  A missing
    model leaves no error behind.  Each case runs in a fresh process: loading happens once."""
    assert os.path.exists(SHIPPED_MODEL)
    e = dict(os.environ)
    e.pop("CONVGEMM_COSTMODEL", None)
    if env is not None:
        e["CONVGEMM_COSTMODEL"] = env
    code = ("import convgemm_amd as cg; L = cg.lib(); "
            "print(cg.get_planner(), cg." + expect + ", L.convgemm_get_last_error())")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=e, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    got, want, err = r.stdout.split()[-3:]
    assert got == want and err == "0", r.stdout


def test_shipped_cost_model_file_loads():
    """The shipped model has this build's feature count and file format."""
    cg.costmodel_load(SHIPPED_MODEL)
    with open(SHIPPED_MODEL, "rb") as f:
        assert f.read(8) == b"CGMLP004"


def test_tuning_database_holds_full_candidate_lists(tmp_path):
    """A database line with CONVGEMM_MAX_CANDIDATES timings (over 16 KB of text, with
    the 1e30 placeholder the save writes for failed candidates) loads whole."""
    d = cg.make_desc(cg.FWD, cg.NHWC, 2, 8, 8, 16, 16, 3, 3, 1, 1, 1, 1)
    key = "0 0 1 2 8 8 16 16 3 3 1 1 1 1 1 1"  # F32, forward NHWC, 16-byte loads, then the shape
    n = cg.MAX_CANDIDATES
    us = [1e30 if i % 4 == 1 else 1000.0 + i / 7 for i in range(n)]
    line = key + f" {n} " + " ".join(f"{i % 31} {1 + i % 200} {u:.3f}" for i, u in enumerate(us))
    assert len(line) > 1 << 14
    path = tmp_path / "long.tunedb"
    path.write_text("# header\n" + line + "\n")
    cg.tune_clear()
    try:
        cg.tune_load(str(path))
        got = cg.tune_get(d, cg.F32)
    finally:
        cg.tune_clear()
    assert len(got) == n
    assert [(c, s) for c, s, _ in got] == [(i % 31, 1 + i % 200) for i in range(n)]
    assert all(abs(g - u) <= 1e-3 * u for (_, _, g), u in zip(got, us))


def test_tuning_database_renumbers_configurations(tmp_path):
    """Configuration ids are tile shape + raster order x (number of shapes); a database
    written by a build with fewer tile shapes (its header records them) is renumbered on
    load, and a save records this build's counts so that a reload is the identity."""
    d = cg.make_desc(cg.FWD, cg.NHWC, 2, 8, 8, 16, 16, 3, 3, 1, 1, 1, 1)
    key = "0 0 1 2 8 8 16 16 3 3 1 1 1 1 1 1"
    ns = cg.num_tile_shapes(cg.F32)
    old = ns - 6  # a writer with six fewer FP32 shapes
    old_ids = [3, old + 5, 2 * old + 1]  # shape 3 order 0, shape 5 order 1, shape 1 order 2
    line = key + " 3 " + " ".join(f"{c} 1 {10.0 + i:.3f}" for i, c in enumerate(old_ids))
    path = tmp_path / "old.tunedb"
    path.write_text(f"# header\n# tile shapes (f32 bf16 fp8): {old} 40 30\n{line}\n")
    cg.tune_clear()
    try:
        cg.tune_load(str(path))
        got = [c for c, _, _ in cg.tune_get(d, cg.F32)]
        assert got == [3, ns + 5, 2 * ns + 1], got
        again = tmp_path / "again.tunedb"
        cg.tune_save(str(again))
        assert f"# tile shapes (f32 bf16 fp8): {ns} {cg.num_tile_shapes(cg.BF16)} " in again.read_text()
        cg.tune_clear()
        cg.tune_load(str(again))
        assert [c for c, _, _ in cg.tune_get(d, cg.F32)] == got
    finally:
        cg.tune_clear()


def test_backward_data_forward_form_has_its_own_key(tmp_path):
    """Backward data of a unit-stride NHWC convolution runs on the forward kernels
    (the forward form: dY convolved with the kernel turned by 180 degrees, c and kn
    swapped, padding (k - 1) * dil - pad).  Its timings include reading dX (dX += ...),
    so they are filed under the backward-data description, never under the forward
    problem with the same GEMM: a database holding both serves each its own plan."""
    fwd = cg.make_desc(cg.FWD, cg.NHWC, 2, 8, 8, 24, 16, 3, 3, 1, 1, 1, 1)   # c 24 -> kn 16
    dgr = cg.make_desc(cg.DGRAD, cg.NHWC, 2, 8, 8, 16, 24, 3, 3, 1, 1, 1, 1)  # its forward form: c 24, kn 16
    fkey = "0 0 1 2 8 8 24 16 3 3 1 1 1 1 1 1"
    dkey = "0 4 1 2 8 8 16 24 3 3 1 1 1 1 1 1"
    ns = cg.num_tile_shapes(cg.F32)
    path = tmp_path / "both.tunedb"
    path.write_text(f"# tile shapes (f32 bf16 fp8): {ns} {cg.num_tile_shapes(cg.BF16)} "
                    f"{cg.num_tile_shapes(cg.FP8_E4M3)}\n"
                    f"{fkey} 1 3 1 10.000\n{dkey} 1 6 2 20.000\n")
    cg.tune_clear()
    try:
        cg.tune_load(str(path))
        assert [(c, s) for c, s, _ in cg.tune_get(fwd, cg.F32)] == [(3, 1)]
        assert [(c, s) for c, s, _ in cg.tune_get(dgr, cg.F32)] == [(6, 2)]
        assert cg.plan(fwd)[0] == 3 and cg.plan(dgr) == (6, 2)
    finally:
        cg.tune_clear()


def _taps_brute(r, pad, k, dil, s):
    # input rows i = r + s*t receive kernel tap kx when i + pad - kx*dil is a multiple of s
    return sum(1 for kx in range(k) if (r + pad - kx * dil) % s == 0)


@pytest.mark.parametrize("k,pad,st,dil", [(3, 1, 2, 1), (1, 0, 2, 1), (5, 2, 2, 1), (3, 1, 3, 1), (3, 0, 2, 2),
                                          (4, 1, 2, 1), (7, 3, 4, 1)])
def test_dgrad_class_order_most_taps_first(k, pad, st, dil):
    """Backward data launches the stride residue classes with the most kernel taps first
    (a stable order: ties keep class order); every class appears once."""
    d = cg.make_desc(cg.DGRAD, cg.NHWC, 2, 16, 16, 8, 8, k, k, pad, pad, st, st, dil, dil)
    order, taps = cg.dgrad_class_order(d)
    n = st * st
    if n > 16:
        assert order == []
        return
    assert sorted(order) == list(range(n))
    want = [_taps_brute(c // st, pad, k, dil, st) * _taps_brute(c % st, pad, k, dil, st) for c in range(n)]
    assert taps == want
    assert [taps[c] for c in order] == sorted(want, reverse=True)
    for i in range(1, n):  # stable among equal tap counts
        if taps[order[i]] == taps[order[i - 1]]:
            assert order[i] > order[i - 1]


def test_dgrad_class_order_3x3_stride2():
    """3x3, stride 2, pad 1: classes (0,0) 1 tap, (0,1) and (1,0) 2 taps, (1,1) 4 taps."""
    d = cg.make_desc(cg.DGRAD, cg.NHWC, 1, 8, 8, 4, 4, 3, 3, 1, 1, 2, 2)
    assert cg.dgrad_class_order(d) == ([3, 1, 2, 0], [1, 2, 2, 4])
