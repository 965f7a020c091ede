"""AddressSanitizer + UndefinedBehaviorSanitizer build of the host C++: the FP64 oracle
(oracle/mwref.cpp) and the product's host numerics (mrchem_amd/csrc/mwg_core.cpp), driven
by tests/cxx/sanitize_driver.cpp.  CPU only; the sanitizers abort on the first error."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    exe = tmp_path / "sanitize_driver"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-fopenmp", "-I", os.path.join(ROOT, "mrchem_amd", "csrc"),
           os.path.join(ROOT, "tests", "cxx", "sanitize_driver.cpp"), os.path.join(ROOT, "oracle", "mwref.cpp"),
           os.path.join(ROOT, "mrchem_amd", "csrc", "mwg_core.cpp"), "-o", str(exe)]
    subprocess.run(cmd, check=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "sanitized host run: ok" in r.stdout
