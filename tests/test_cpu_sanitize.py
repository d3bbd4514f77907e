"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer.

io.cpp (checkpoint loading), mesh.cpp (rest-shape construction) and
init.cpp (particle placement) are compiled with g++ and
-fsanitize=address,undefined together with tests/sanitize/host_fuzz.cpp,
which feeds them a valid checkpoint built from the reference's packed
particles, every truncation of its rank file and header, randomly garbled
copies, broken meshes (out-of-range, negative, repeated indices,
non-manifold edges, NaN vertices) and empty, huge and NaN placement
requests.  Every call must return a status with no sanitizer report and no
exception escaping the C ABI.  No GPU call is made (the HIP runtime is only
linked).  Found and fixed with it: a checkpoint header naming a directory
threw through bl_ibm_load_checkpoint, and a box of more cells than an int
holds overflowed bl_positions_create's search grid.
"""
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT

CSRC = os.path.join(ROOT, "biblbm_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.exists("/opt/rocm/lib/libamdhip64.so"),
                    reason="needs g++ and the HIP runtime library")
def test_host_code_under_sanitizers(tmp_path):
    lib = os.path.join(ROOT, "biblbm_amd", "libbiblbm.so")
    if not os.path.exists(lib):
        pytest.skip("libbiblbm.so not built")
    exe = str(tmp_path / "host_fuzz")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + CSRC,
           "-I" + os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "sanitize", "host_fuzz.cpp")]
    cmd += [os.path.join(CSRC, f) for f in ("io.cpp", "mesh.cpp", "init.cpp", "errors.cpp")]
    cmd += ["-L" + os.path.dirname(lib), "-lbiblbm", "-L/opt/rocm/lib", "-lamdhip64",
            "-Wl,-rpath," + os.path.dirname(lib), "-Wl,-rpath,/opt/rocm/lib", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, GOLDEN, str(tmp_path)], capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0
    assert "no sanitizer report" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
