"""The library's host C++ (natural-neighbour interpolation and NaN fill,
csrc/natural.cpp) and the FP64 oracle (oracle/aldic_oracle.c) built for the
CPU with AddressSanitizer and UndefinedBehaviorSanitizer and run on seeded
inputs and edge cases (tests/sanitize/).  The HIP part of the library is not
in this build: its host code is the C ABI around kernel launches."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None,
                    reason="needs g++ and make")
def test_host_code_under_asan_and_ubsan(tmp_path):
    r = subprocess.run(["make", "-s", "-C", os.path.join(HERE, "sanitize"), f"OUT={tmp_path}"],
                       capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0
    assert "every check held" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
