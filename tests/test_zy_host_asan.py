"""The library's host code under AddressSanitizer: tests/native/asan_driver (built by
tests/native/Makefile from the library's sources with the host side instrumented)
drives every entry point of the C-ABI -- both precision forms and methods, the three
boundary forms, injection, compaction, re-sorts, graphs, the no-allocation step, a
batch with its graph and the autotuner, the adjoint, refused calls and every destroy
order.  Any invalid host memory access fails the run; the second run also checks for
leaks, with only the ROCm runtime's own allocations suppressed."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "native", "asan_driver")
SUPP = os.path.join(HERE, "native", "lsan.supp")


@pytest.mark.parametrize("leaks", [False, True])
def test_host_code_under_address_sanitizer(leaks):
    assert os.path.exists(DRIVER), "tests/native/asan_driver is not built (make -C tests/native)"
    env = dict(os.environ)
    # A quarantine large enough that nothing freed is recycled while the HIP runtime tears
    # itself down at exit: recycling a chunk of the runtime's pinned host memory after the
    # runtime has unloaded trips a CHECK inside the sanitizer's own allocator (the default
    # 256 MB fills up once the driver's larger clouds have been freed).  A larger
    # quarantine keeps freed memory poisoned for longer.
    env["ASAN_OPTIONS"] = ("detect_leaks=1" if leaks else "detect_leaks=0") + ":quarantine_size_mb=2048"
    if leaks:
        env["LSAN_OPTIONS"] = f"suppressions={SUPP}"
    r = subprocess.run([DRIVER], env=env, capture_output=True, text=True, timeout=240)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "LeakSanitizer" not in out, out[-4000:]
    assert r.returncode == 0 and "asan_driver: ok" in r.stdout, out[-4000:]
