"""The C++ layer (include/biblbm_plb.h) compiles as C++17 with the example
progs written against it; the progs' binaries are built and linked against
the library.  CPU only (nothing is run)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


import pytest

EXAMPLES = ["capsule_in_shear", "rbc_in_pipe"]


@pytest.mark.parametrize("name", EXAMPLES)
def test_facade_and_example_compile(name):
    src = os.path.join(ROOT, "biblbm_amd", "examples", name + ".cpp")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I" + os.path.join(ROOT, "include"), src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("name", EXAMPLES)
def test_example_binary_links_the_library(name):
    exe = os.path.join(ROOT, "biblbm_amd", "examples", name)
    assert os.path.exists(exe), "built by __graft_entry__.build()"
    r = subprocess.run(["ldd", exe], capture_output=True, text=True)
    assert "libbiblbm.so" in r.stdout and "not found" not in r.stdout.split("libbiblbm.so")[1].splitlines()[0]
