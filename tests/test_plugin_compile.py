"""The VSIPL++ plugin (include/ovxhip/vsip_plugin.hpp) compiles against the
reference's headers and dispatches the reference's own expression types to
be::user: tests/cpp/plugin_instantiate.cpp instantiates every evaluator
(Fftm c->c / r->c / c->r and Fft factories, the three fused fastconv forms,
the per-row vector form, vmmul and the by-value Fftm on Resident blocks,
Correlation, Convolution with its generic fallback) and static_asserts that
be::user takes each expression it should (and not the Dense-block vmmul).

OpenVSIP's headers include the configure-generated ovxx/detail/config.hpp.
The test writes a small one into a temporary directory; it is a compile
fixture only (it selects no backend, and nothing built with it runs or is
compared with anything).  Compiled with the threading macro off and on
(src/ovxx/thread.hpp:16-23 rewrites the thread_local keyword either way),
in C++11 and C++17, and with split complex storage.  Skipped where the
reference tree is absent."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"
TU = os.path.join(ROOT, "tests", "cpp", "plugin_instantiate.cpp")

CONFIG = """\
/* compile fixture for tests/test_plugin_compile.py */
#define OVXX_HAS_EXCEPTIONS 1
#define OVXX_ALLOC_ALIGNMENT 32
#define OVXX_DEFAULT_COMPLEX_STORAGE_SPLIT {split}
#define OVXX_TIMER_SYSTEM 1
#define OVXX_DFT_FFT 1
#define VSIP_BIG_ENDIAN 0
#define HAVE_STD_ISFINITE 1
#define HAVE_STD_ISNAN 1
#define HAVE_STD_ISNORMAL 1
"""

pytestmark = [
    pytest.mark.skipif(not os.path.isdir(os.path.join(REF_SRC, "ovxx")), reason="reference tree absent"),
    pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available"),
]


def _compile(tmp_path, std, threading, split):
    cfg = tmp_path / "ovxx" / "detail"
    cfg.mkdir(parents=True, exist_ok=True)
    (cfg / "config.hpp").write_text(CONFIG.format(split=split))
    cmd = ["g++", f"-std={std}", "-fsyntax-only", "-Wall", "-Wno-terminate", "-I", str(tmp_path), "-I", REF_SRC,
           "-I", os.path.join(ROOT, "include"), TU]
    if threading:
        cmd.insert(1, "-DOVXX_ENABLE_THREADING=1")
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-6000:]
    # no diagnostic may point into this repository's headers
    ours = [ln for ln in out.splitlines() if "include/ovxhip" in ln and ("warning" in ln or "error" in ln)]
    assert not ours, "\n".join(ours)


@pytest.mark.parametrize("threading", [0, 1])
def test_plugin_compiles_and_dispatches(tmp_path, threading):
    _compile(tmp_path, "c++11", threading, 0)


def test_plugin_compiles_cxx17(tmp_path):
    _compile(tmp_path, "c++17", 1, 0)


def test_plugin_compiles_split_storage(tmp_path):
    _compile(tmp_path, "c++11", 0, 1)
