"""The host layer under AddressSanitizer and UndefinedBehaviorSanitizer: the
host sources of libaccint_host (csrc/host/*.cpp) and the driver
tests/sanitize/host_driver.cpp are compiled with g++ and
-fsanitize=address,undefined into a temporary directory and run (no device is
used; the driver exercises the JSON readers, BasisSet, the Fock builders and
their threading strategies, the dispatch table and the no-device error
paths).  Any sanitizer report fails the test."""
import glob
import os
import shutil
import subprocess
import tempfile

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "libaccint-public_amd")
LIBDIR = os.path.join(PKG, "libaccint")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
@pytest.mark.skipif(not os.path.exists(os.path.join(LIBDIR, "liblai_device.so")), reason="device library not built")
def test_host_layer_under_asan_and_ubsan():
    srcs = sorted(glob.glob(os.path.join(PKG, "csrc", "host", "*.cpp")))
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "host_driver")
        cmd = ["g++", "-std=c++20", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
               "-fno-sanitize-recover=undefined", "-I" + os.path.join(REPO, "include"),
               os.path.join(REPO, "tests", "sanitize", "host_driver.cpp")] + srcs + \
              ["-L" + LIBDIR, "-llai_device", "-Wl,-rpath," + LIBDIR, "-pthread", "-o", exe]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        assert r.returncode == 0, r.stdout[-4000:]
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
                   UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
                   LIBACCINT_DATA_DIR=os.path.join(PKG, "share", "basis_sets"), HIP_VISIBLE_DEVICES="")
        env.pop("LIBACCINT_MAX_FOCK_THREADLOCAL_BYTES", None)
        r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env, timeout=300)
        assert r.returncode == 0 and "host driver ok" in r.stdout, r.stdout[-4000:]
        assert "runtime error" not in r.stdout and "AddressSanitizer" not in r.stdout, r.stdout[-4000:]
