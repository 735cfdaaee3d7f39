"""The library's host code (planner, tree executor, grouped launches, request loop, memory
pool) under AddressSanitizer and UndefinedBehaviorSanitizer: tests/native/capi_check, built
with both on the host side only (make -C tests/native), drives the C ABI on the GPU over flat
binary, unary, n-ary and permute calls, the allocation-free entry, plain and sliced trees,
the request loop with an injected status, and the grouped launches, each checked against a
host loop.  A sanitizer finding aborts the program and fails the test."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "_build", "capi_check")


def test_capi_under_host_sanitizers():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    assert os.path.exists(EXE), "build it first: make -C tests/native"
    env = dict(os.environ)
    # the HIP runtime keeps allocations until exit; leaks are not what this run looks for
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-4000:])
    assert r.stdout.strip().startswith("ok "), r.stdout + r.stderr
