"""The C++ drop-in header (include/fvk_dropin.hpp) compiles against the
reference's headers with warnings as errors (where the reference tree is
present; the GPU run of the compiled program is tests/test_gpu_dropin.py)."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INCLUDE = "/root/reference/include"


@pytest.mark.skipif(not os.path.isdir(REF_INCLUDE) or shutil.which("g++") is None,
                    reason="reference headers or g++ not present")
def test_dropin_header_compiles_against_reference_headers():
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-Wno-unused-parameter",
           f"-I{REF_INCLUDE}", f"-I{os.path.join(REPO, 'include')}", "-x", "c++",
           os.path.join(REPO, "tests", "dropin_check.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
