"""The host code that reads untrusted input, and the serial restatement,
built with AddressSanitizer and UndefinedBehaviorSanitizer and run on the
CPU (tests/sanitize/): the YAML parameter loader on reference-style,
malformed and random files; the checkpoint record check on 20,000
truncated, extended, bit-flipped, resized and random records; and the
restatement through frames with map shifts, non-finite points, brute-force
and threaded kernel inference, several LiDARs and the DS-PHD/MIB method.
Any sanitizer report aborts the driver."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None, reason="needs g++ and make")
def test_host_code_under_asan_and_ubsan(tmp_path):
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "sanitize")], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="4")
    r = subprocess.run([os.path.join(HERE, "sanitize", "_build", "san_driver"), str(tmp_path)], env=env,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and "sanitize ok" in r.stdout, (r.stdout[-2000:], r.stderr[-4000:])
