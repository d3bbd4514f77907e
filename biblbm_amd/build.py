"""Compiles the gfx950 library biblbm_amd/libbiblbm.so in-tree with hipcc.

Every call recompiles every source (the build check must exercise the
compiler each time); objects go to biblbm_amd/_build/, the shared library to
biblbm_amd/libbiblbm.so.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbiblbm.so")

SOURCES = ["errors.cpp", "lattice.hip", "ibm.hip", "mesh.cpp", "ensemble.hip", "walls.hip", "interactions.hip", "io.cpp",
           "init.cpp", "tune.cpp", "rigid.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions stay off (the compiler may otherwise form
    # v_pk_* FP32 ops from scalar or float2 code)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # a*b+c is fused within one source expression only: the rounding of a
    # cell's arithmetic then does not depend on the kernel it is inlined in
    # (every launch shape and block order gives bitwise the same results)
    "-ffp-contract=on",
    "-Wno-unused-command-line-argument", "-Wno-unused-result",
]


def _hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def compile_one(src):
        obj = os.path.join(BUILD, src + ".o")
        cmd = [hipcc] + HIPCC_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [hipcc, "-O3", "-fPIC", "-std=c++17", "-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
        if verbose and r.stderr.strip():
            sys.stderr.write(r.stderr)
        return obj

    with ThreadPoolExecutor(max_workers=jobs or min(8, len(srcs))) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    build_examples()


EXAMPLES = ["capsule_in_shear", "rbc_in_pipe", "initial_condition"]
INCLUDE = os.path.join(os.path.dirname(HERE), "include")


def build_examples():
    """C++ progs over include/biblbm_plb.h, linked against the library
    (found next to the binary's folder at run time)."""
    for name in EXAMPLES:
        src = os.path.join(HERE, "examples", name + ".cpp")
        out = os.path.join(HERE, "examples", name)
        rocm_lib = "/opt/rocm/lib"
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I" + INCLUDE, src, "-L" + HERE, "-lbiblbm", "-L" + rocm_lib,
               "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath," + rocm_lib, "-o", out]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("example build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return LIB


# The reference's own progs, compiled unchanged against the Palabos-named
# headers in include/ (palabos3D.h, fsi/headers.h, ...) and linked against
# the library, when the reference checkout is present.  Outputs go to
# examples/ref_progs/ (not versioned).
REF_PROGS = ["capsule_in_shear/capsule_in_shear.cpp", "rbc_in_pipe/rbc_in_pipe.cpp", "rbc_in_shear/rbc_in_shear.cpp",
             "rbcs_in_shear/rbcs_in_shear.cpp"]
REF_PROG_FLAGS = ["-std=gnu++14", "-fpermissive", "-w", "-O2"]


def ref_prog_command(src, out):
    rocm_lib = "/opt/rocm/lib"
    return (["g++"] + REF_PROG_FLAGS + ["-I" + INCLUDE, src, "-L" + HERE, "-lbiblbm", "-L" + rocm_lib,
                                        "-Wl,-rpath,$ORIGIN/../..", "-Wl,-rpath," + rocm_lib, "-o", out])


def build_ref_progs(ref="/root/reference", jobs=4):
    progs = os.path.join(ref, "progs")
    if not os.path.isdir(progs):
        return []
    outdir = os.path.join(HERE, "examples", "ref_progs")
    os.makedirs(outdir, exist_ok=True)

    def one(rel):
        out = os.path.join(outdir, os.path.basename(os.path.dirname(rel)))
        r = subprocess.run(ref_prog_command(os.path.join(progs, rel), out), capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("reference prog %s does not build against include/: %s" % (rel, r.stderr[-4000:]))
        return out

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        return list(ex.map(one, REF_PROGS))


if __name__ == "__main__":
    print(build(verbose=True))
