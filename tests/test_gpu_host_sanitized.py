"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/host_sanitize/_bin/abi_driver holds the library's sources with their
host code built with -fsanitize=address,undefined (device code as usual) and a
driver that walks the C ABI end to end: both precisions with several types,
exclusions, FENE, tabulated pair and bond tables, Langevin runs, energies,
pairs, the RDF, interaction removal, the allocation-free energy, checkpoint
state, pruned lists, the learned force with training and Adam, the batched
replicas with the autotuner, and error paths.  Any invalid host access, leak of
the library's own allocations (the HIP and HSA runtimes' are suppressed) or
undefined arithmetic aborts it.  The driver is built on the CPU by
tests/host_sanitize/build_driver.py; the test is skipped, saying so, when the
binary is missing or older than the sources.
"""
import glob
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SAN = os.path.join(HERE, "host_sanitize")
DRIVER = os.path.join(SAN, "_bin", "abi_driver")


def test_host_code_runs_clean_under_sanitizers(gpu):
    if not os.path.exists(DRIVER):
        pytest.skip("the sanitized driver is not built (python tests/host_sanitize/build_driver.py)")
    sources = glob.glob(os.path.join(ROOT, "espressopp_amd", "csrc", "*")) + \
        glob.glob(os.path.join(ROOT, "include", "*.h")) + [os.path.join(SAN, "abi_driver.cpp")]
    if os.path.getmtime(DRIVER) < max(os.path.getmtime(p) for p in sources):
        pytest.skip("the sanitized driver is older than the sources; rebuild it")
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               LSAN_OPTIONS="suppressions=" + os.path.join(SAN, "lsan.supp"),
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([DRIVER], capture_output=True, text=True, env=env, timeout=600)
    print(r.stdout[-2000:], r.stderr[-6000:], flush=True)
    assert r.returncode == 0, r.stderr[-6000:]
    assert "host sanitize driver: clean" in r.stdout
