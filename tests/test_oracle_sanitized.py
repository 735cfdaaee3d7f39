"""The C oracle built with AddressSanitizer and UndefinedBehaviorSanitizer (gcc) and run over
its known-answer driver (oracle/sanitize_check.c): every element type, empty dimensions,
scalar and embedded outputs, three operands, rank 12, sx / sy scaling and the argument
checks.  A sanitizer finding aborts the driver, so the test fails on it."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None,
                    reason="needs gcc and make")
def test_oracle_under_address_and_undefined_behavior_sanitizers():
    r = subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().startswith("ok "), r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
