"""Host sanitizer run of the oracle (CPU): oracle/san_driver.c drives every
stage of the restatement (box, keys, presort with coincident particles and the
duplicate fix-up, tree, tiles, group walk, lists, adjoint, direct sum, lattice
coefficients, mirror walk, far field with each extrinsic correction) in a
build with AddressSanitizer and UndefinedBehaviorSanitizer, every report
fatal.  The driver's own consistency checks must pass as well."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_under_asan_ubsan():
    if shutil.which(os.environ.get("CC", "cc")) is None:
        pytest.skip("no C compiler")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="2")
    r = subprocess.run([os.path.join(ROOT, "oracle", "_san", "san_driver")], env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "san ok" in r.stdout
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
