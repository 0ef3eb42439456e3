"""The C++ host mirror (include/wrach_api.hpp): builds everywhere, runs the
reference's API-level tests on a GPU."""
import os
import subprocess

import pytest

from tests.conftest import ROOT

SRC = os.path.join(ROOT, "tests", "cpp", "test_wrach_api.cpp")


def build_cpp_test(lib_path, out_dir):
    exe = os.path.join(out_dir, "test_wrach_api")
    libdir = os.path.dirname(lib_path)
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"), SRC,
                    "-o", exe, f"-L{libdir}", "-l:libwrach_hip.so", f"-Wl,-rpath,{libdir}"],
                   check=True)
    return exe


def test_cpp_mirror_builds(wrach_lib_path, tmp_path):
    assert os.path.exists(build_cpp_test(wrach_lib_path, str(tmp_path)))


@pytest.mark.gpu
def test_cpp_mirror_runs_reference_api_tests(wrach_lib_path, tmp_path):
    exe = build_cpp_test(wrach_lib_path, str(tmp_path))
    print("[cpp] running", exe, flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "ok (0 failures)" in r.stdout
