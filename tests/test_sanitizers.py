"""The host layer under AddressSanitizer and UndefinedBehaviorSanitizer (no GPU).

tests/test_abi.py drives the C host layer through its validation, quick
returns, planner start-up, allocation-failure paths and exports without a
GPU.  Here it runs again, in a fresh process, against a build whose host C is
compiled with gcc's -fsanitize=address,undefined (convgemm_amd.build.
build_sanitized; the kernels are the normal objects), with gcc's ASan runtime
preloaded: any invalid access, leak-free misuse or undefined behaviour in the
host code ends the run with a sanitizer report.
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_layer_under_asan_and_ubsan(tmp_path):
    from convgemm_amd import build as cgbuild
    lib = cgbuild.build_sanitized(str(tmp_path / "asan"))
    undef = subprocess.run(["nm", "-D", "--undefined-only", lib], stdout=subprocess.PIPE, text=True).stdout
    assert "__asan_report_load" in undef and "__ubsan_handle" in undef  # the host code is instrumented
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], stdout=subprocess.PIPE, text=True).stdout.strip()
    env = dict(os.environ, CONVGEMM_LIB=lib, LD_PRELOAD=os.path.realpath(asan),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",  # the interpreter itself is not leak-clean
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_abi.py")], cwd=ROOT, env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=900)
    out = r.stdout
    assert "AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
    assert " passed" in out, out[-4000:]
