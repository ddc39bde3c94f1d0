"""The C++ adapter (include/gpgmp_hip_solver.hpp) compiles against the C ABI, and a
host loop written like DiffusionModel::oclSolve through it matches the fused device
loop bit for bit."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "host_loop.cpp")
EXE = os.path.join(ROOT, "tests", "cpp", "host_loop")


def _build():
    lib_dir = os.path.join(ROOT, "inchman_amd")
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include",
           "-D__HIP_PLATFORM_AMD__", SRC, "-L", lib_dir, "-lgpgmp_hip", "-Wl,-rpath," + lib_dir, "-o", EXE]
    subprocess.run(cmd, check=True, capture_output=True, text=True)


def test_adapter_compiles():
    _build()
    assert os.path.exists(EXE)


@pytest.mark.gpu
def test_adapter_host_loop_matches_fused(gpu):
    if not os.path.exists(EXE):
        _build()
    r = subprocess.run([EXE, "300"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
