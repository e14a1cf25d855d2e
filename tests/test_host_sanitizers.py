"""The host code of the library under AddressSanitizer and UndefinedBehaviorSanitizer: the CPU
tests that drive it (basis, network handles and their decode checks, pool validation, tuning,
error paths, the ABI) rerun in a subprocess against libhipotf_asan.so (binsim_amd/build.py
--asan: host code sanitized, device code as usual) with the compiler's ASan runtime preloaded.
Any sanitizer report fails the test.  No GPU is used."""
import os
import subprocess
import sys

import pytest

from binsim_amd import build as BLD

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_TESTS = ["tests/test_abi.py", "tests/test_network_decode.py", "tests/test_autotune.py",
              "tests/test_colex.py", "tests/test_oracle_spin.py"]


@pytest.mark.skipif(not os.path.exists(BLD.LIB_ASAN), reason="libhipotf_asan.so not built "
                    "(python -m binsim_amd.build --asan)")
@pytest.mark.skipif(BLD.asan_runtime() is None, reason="no ASan runtime for the compiler")
def test_host_paths_clean_under_asan_ubsan():
    env = dict(os.environ, LD_PRELOAD=BLD.asan_runtime(), HIPOTF_LIBRARY=BLD.LIB_ASAN,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-m", "not gpu", "-p", "no:cacheprovider",
                        *HOST_TESTS], cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
    assert " passed" in out
