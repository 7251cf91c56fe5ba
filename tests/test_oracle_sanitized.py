"""The FP64 oracle under AddressSanitizer and UndefinedBehaviorSanitizer (CPU):
oracle/sanitize_driver.c runs an LJ fluid with Langevin, bonded WCA + FENE
chains with exclusions and tabulated pair and bond tables through the oracle;
any invalid access, leak or undefined arithmetic aborts it."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_runs_clean_under_sanitizers():
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", OMP_NUM_THREADS="2")
    r = subprocess.run([os.path.join(ORACLE, "_ref", "sanitize_driver")], capture_output=True, text=True, env=env,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "sanitize driver: clean" in r.stdout
