"""The C++ layer (include/biblbm_plb.h) compiles as C++17 with the example
prog written against it; the prog's binary is built and linked against the
library.  CPU only (nothing is run)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_facade_and_example_compile():
    src = os.path.join(ROOT, "biblbm_amd", "examples", "capsule_in_shear.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I" + os.path.join(ROOT, "include"), src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_example_binary_links_the_library():
    exe = os.path.join(ROOT, "biblbm_amd", "examples", "capsule_in_shear")
    assert os.path.exists(exe), "built by __graft_entry__.build()"
    r = subprocess.run(["ldd", exe], capture_output=True, text=True)
    assert "libbiblbm.so" in r.stdout and "not found" not in r.stdout.split("libbiblbm.so")[1].splitlines()[0]
