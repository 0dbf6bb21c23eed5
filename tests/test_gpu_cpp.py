"""The reference's hot-path unit tests restated in C++ against the YAFEL-named
header layer (tests/cpp/test_hot_path.cpp, built by __graft_entry__.build())."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_hot_path_driver():
    exe = os.path.join(ROOT, "tests", "cpp", "test_hot_path")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/test_hot_path is not built: run __graft_entry__.build()")
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=100)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
