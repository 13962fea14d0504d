"""The library's host code under AddressSanitizer and
UndefinedBehaviorSanitizer: the cloud generator (k-d tree kNN, visibility
and balance passes), the partGrid reader and writer, normals, sub-stencils,
RCM, the restart and output files and the Hicks-Henne fields, built from
their sources with g++ and driven by tests/host_sanitize.cpp.  CPU only."""
import os
import shutil
import subprocess

import pytest

from conftest import REPO

CSRC = os.path.join(REPO, "c-_serial_amd", "csrc")
HOST_SOURCES = ["cloud.cpp", "setup.cpp", "restart.cpp", "output.cpp", "shape.cpp"]


@pytest.mark.timeout(600)
def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ absent")
    exe = tmp_path / "host_sanitize"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fopenmp",
           "-ffp-contract=off", "-I" + os.path.join(REPO, "include"), "-I" + CSRC,
           os.path.join(REPO, "tests", "host_sanitize.cpp"),
           *[os.path.join(CSRC, s) for s in HOST_SOURCES], "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="4")
    proc = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, env=env,
                          timeout=500)
    assert proc.returncode == 0, proc.stdout[-2000:] + proc.stderr[-4000:]
    assert "host code OK" in proc.stdout
    assert "runtime error" not in proc.stderr and "AddressSanitizer" not in proc.stderr
