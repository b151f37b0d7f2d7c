"""The C ABI's CPU tests (tests/test_abi_cpu.py: every validation and
refusal path on a validation-only context) rerun against a host-sanitized
build of the library: the ABI's host sources compiled with AddressSanitizer
and UndefinedBehaviorSanitizer (`make -C asc-cpp_amd sanitize`, part of the
default build, output in asc-cpp_amd/build/san/).  The GCC sanitizer runtimes
are preloaded into a child interpreter; a report aborts it.  Leak detection
is off (the interpreter itself is not instrumented)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "asc-cpp_amd", "build", "san", "libasc_sparse_hip.so")


def runtime(name):
    p = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True)
    path = p.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def test_abi_cpu_tests_under_address_and_ub_sanitizers():
    if not os.path.exists(LIB):
        pytest.skip("the sanitizer build is made by __graft_entry__.build() (make sanitize)")
    asan, ubsan = runtime("libasan.so"), runtime("libubsan.so")
    if asan is None or ubsan is None:
        pytest.skip("no GCC sanitizer runtimes on this machine")
    syms = subprocess.run(["nm", "-D", LIB], capture_output=True, text=True).stdout
    assert "__asan_report_load" in syms and "__ubsan_handle" in syms, "build is not instrumented"
    env = dict(os.environ, LD_PRELOAD=f"{asan}:{ubsan}", ASC_SPARSE_HIP_LIBRARY=LIB,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    p = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_abi_cpu.py")],
                       capture_output=True, text=True, env=env, cwd=ROOT, timeout=600)
    print(p.stdout[-3000:])
    print(p.stderr[-3000:])
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    assert "passed" in p.stdout and "ERROR: AddressSanitizer" not in p.stderr
    assert "runtime error" not in p.stderr  # UBSan's report line
