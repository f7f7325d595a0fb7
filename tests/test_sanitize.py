"""The host code of the C-ABI under AddressSanitizer and UndefinedBehaviorSanitizer
(no GPU needed): the library is rebuilt with both sanitizers on its host
compilation and tests/sanitize/host_driver.cpp calls every entry point that
runs without a device (generators, tables, time control, parameter defaults)
and the argument and handle checks that return before any device call.  Any
sanitizer report, leak included, fails the run."""
import os
import subprocess

from elsim_amd import build


def test_host_code_is_clean_under_asan_and_ubsan():
    drv = build.build_sanitized()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=23",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([drv], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "ALL OK" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
