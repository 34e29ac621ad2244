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


@pytest.mark.gpu
@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_cpp_host_api_on_gpu(tmp_path):
    """The C++ header end to end on the device, as the VSIPL++ plugin uses
    it: host pointers, split storage, the fastconv plan cache and Storage
    (tests/cpp/host_api_gpu.cpp, checked against an FP64 DFT)."""
    if not os.path.exists(os.path.join(LIB_DIR, "libovxhip.so")):
        pytest.fail("libovxhip.so not built")
    exe = str(tmp_path / "host_api_gpu")
    src = os.path.join(ROOT, "tests", "cpp", "host_api_gpu.cpp")
    subprocess.run(["g++", "-std=c++17", "-O2", "-pthread", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), src,
                    "-L", LIB_DIR, "-lovxhip", "-Wl,-rpath," + LIB_DIR, "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")
