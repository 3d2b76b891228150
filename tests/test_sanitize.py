"""The CPU oracle and the C ABI's host-side code run clean under
AddressSanitizer and UndefinedBehaviorSanitizer (tests/sanitize_build.py
builds both drivers; a rebuild happens only when a source changed). No GPU:
the ABI driver exercises argument checks, host demod, pilot signs and the
tuning overrides, which all return before any device work."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import sanitize_build  # noqa: E402


def _run(binary):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([binary], capture_output=True, text=True, timeout=300, env=env)
    out = r.stdout + r.stderr
    print(out)
    assert r.returncode == 0, out
    assert "AddressSanitizer" not in out and "runtime error" not in out, out
    assert " 0 failed" in out


def test_oracle_under_sanitizers():
    _run(sanitize_build.build_oracle())


@pytest.mark.timeout(600)
def test_c_abi_host_code_under_sanitizers():
    _run(sanitize_build.build_abi())
