"""The C++ host header (include/ovxhip/ovxhip.hpp) compiles against the
C-ABI, links with libovxhip.so and maps statuses to the documented
exceptions (no GPU needed: only argument and length checks run)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(ROOT, "openvsip_amd", "lib")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_cpp_header_compiles_links_and_maps_errors(tmp_path):
    if not os.path.exists(os.path.join(LIB_DIR, "libovxhip.so")):
        pytest.skip("libovxhip.so not built")
    exe = str(tmp_path / "header_check")
    src = os.path.join(ROOT, "tests", "cpp", "header_check.cpp")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), src,
                    "-L", LIB_DIR, "-lovxhip", "-Wl,-rpath," + LIB_DIR, "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")
