"""Served loop admission (host code only, no GPU).

The gate that admits requests into slots against the memory budget
(dygrass_amd/csrc/admission.h) is compiled by itself with g++ together with
tests/native/admission_check.cpp and run under a time limit: a waiter that is
never woken (the deadlock of resident state growing while a request waits)
shows up as a timeout.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
def test_admission_gate(tmp_path):
    exe = str(tmp_path / "admission_check")
    subprocess.run(["g++", "-O1", "-std=c++17", "-pthread", "-Wall", "-Werror",
                    "-I", os.path.join(ROOT, "dygrass_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "admission_check.cpp"), "-o", exe],
                   check=True, capture_output=True, text=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "admission checks passed" in out.stdout
