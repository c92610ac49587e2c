"""Runs the C++ drop-in test program (tests/cpp/test_hmmlike.cpp): the
reference's HMMLike / GLPack gtest cases against the C++ host classes."""
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(__file__), "cpp", "test_hmmlike")


def _bin():
    if not os.path.exists(BIN):
        from glphase_amd import build
        build.build_cpp_tests()
    return BIN


def test_cpp_host_classes_cpu():
    r = subprocess.run([_bin(), "--cpu"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_cpp_hmmlike_on_gpu():
    r = subprocess.run([_bin()], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
