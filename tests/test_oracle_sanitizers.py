"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer (host code
only): oracle/sanitize_main.c runs every primitive's keyGen and evaluation on
exact-size heap buffers over a grid of widths, element sizes and group sizes,
with AES-NI and with the portable AES, and checks each reconstruction
property.  Any out-of-bounds access, leak or undefined operation fails the
run."""
import os
import subprocess
from pathlib import Path

ORACLE = Path(__file__).resolve().parents[1] / "oracle"


def test_oracle_clean_under_asan_ubsan():
    subprocess.run(["make", "-s", "-f", str(ORACLE / "sanitize.mk")], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(ORACLE / "_sanitize" / "fss_sanitize")], env=env, capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "all checks passed" in r.stdout
