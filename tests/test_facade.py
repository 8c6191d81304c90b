"""The C++ interface (include/tfhe_amd.hpp: GPUBinFHEContext, GPUCirBTSContext with the
reference's method names) compiles with a plain host compiler, and its test program
(tests/cpp/facade_test.cpp) passes on the GPU: gate truth tables, batched gates,
EvalFunc / EvalFloor / EvalSign / EvalDecomp, circuit bootstraps from device LWEs in the reference's layout, AES-128 FIPS-197 C.1."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CPP = os.path.join(HERE, "cpp")
EXE = os.path.join(CPP, "facade_test")


def test_cpp_interface_header_compiles():
    src = '#include "tfhe_amd.hpp"\nint main() { return tfhe_amd::GateParams().n == 636 ? 0 : 1; }\n'
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-D__HIP_PLATFORM_AMD__",
                        "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include", "-x", "c++", "-"],
                       input=src, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


def test_cpp_interface_program_links_in_tree_library():
    if not os.path.exists(EXE):
        pytest.skip("test program not built")
    r = subprocess.run(["ldd", EXE], capture_output=True, text=True, timeout=60)
    line = [ln for ln in r.stdout.splitlines() if "libtfhe_amd.so" in ln]
    assert line and "not found" not in line[0], r.stdout
    assert os.path.realpath(line[0].split("=>")[1].split()[0]).startswith(os.path.realpath(ROOT)), line[0]


@pytest.mark.gpu
def test_cpp_interface_program(gpu):
    if not os.path.exists(EXE):
        subprocess.run(["make", "-C", CPP], check=True)
    print("running", EXE, flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().splitlines()[-1] == "facade ok"
