"""The drop-in C++ interface (include/nabo/nabo.h).

CPU: the header compiles with g++ (C++17) for our API example, and libnabo's
own examples/trivial.cpp and examples/usage.cpp compile unchanged against it
(read from a libnabo checkout when one is present, nothing is copied).
GPU: the API example runs through libnabo_hip.so and checks itself.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "api_example.cpp")
BIN = os.path.join(ROOT, "tests", "cpp", "api_example")
LIB = os.path.join(ROOT, "libnabo_amd", "libnabo_hip.so")
REF_EXAMPLES = "/root/reference/examples"


def build_example(out=BIN):
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), SRC, "-o",
           out, LIB, f"-Wl,-rpath,{os.path.dirname(LIB)}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def test_header_compiles(tmp_path):
    build_example(str(tmp_path / "api_example"))


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="no libnabo checkout")
@pytest.mark.parametrize("name", ["trivial.cpp", "usage.cpp"])
def test_reference_examples_compile_unchanged(name):
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                    os.path.join(REF_EXAMPLES, name)], check=True)


@pytest.mark.gpu
def test_api_example_runs_on_gpu():
    exe = BIN if os.path.exists(BIN) else build_example()
    print("case cpp api example", flush=True)
    res = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "API OK" in res.stdout
