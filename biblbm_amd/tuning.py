"""The project's autotuner over the C-ABI (bl_tune_*; the reference's block
decomposition is fixed, src/multiBlock/defaultMultiBlockPolicy3D.h:119-129).

- tune_collide: per AA step parity, the collide-and-stream kernel's threads
  per block, XCD-aware block order and nontemporal hints (even step); the
  lattice state is left as it was.
- tune_coupled: the launch structure of coupled steps (velocity staging
  carried by the collision or not, stages paired on one stream or forked);
  every structure gives the same results, the trials are real steps.
- tune_ibm: the immersed-boundary kernels' launch parameters (workgroup
  sizes of binning, element forces and interpolation, gather lanes per
  vertex, spreading grid), coordinate descent over real steps.
- save / load: the process's tuning table as JSON lines (key: device,
  lattice size, storage, and for coupled steps the vertex count); a lattice
  created after a load starts with the tuned shapes of its size, and
  apply_coupled sets a boundary's tuned structure.  The default file is
  $BIBLBM_TUNE_CACHE or ~/.cache/biblbm_amd/tune.jsonl.
"""
import ctypes
import json
import os

from .capi import call

DEFAULT_PATH = os.environ.get("BIBLBM_TUNE_CACHE",
                              os.path.join(os.path.expanduser("~"), ".cache", "biblbm_amd", "tune.jsonl"))


def tune_collide(lattice):
    """bl_tune_collide: returns the JSON report (trials per parity, best)."""
    buf = ctypes.create_string_buffer(65536)
    call("bl_tune_collide", lattice.handle, buf, len(buf))
    return json.loads(buf.value.decode())


def apply_collide(lattice, best):
    """Sets the per-parity shapes of a tune_collide report's "best"."""
    for parity, key in ((0, "even"), (1, "odd")):
        b = best[key]
        lattice.set_launch_parity(parity, b["block"], b["xcd_remap"], b["nontemporal"])


def tune_coupled(ibm, n_steps=8):
    return ibm.tune(n_steps)


def tune_ibm(ibm, n_steps=4):
    return ibm.tune_launch(n_steps)


def apply_coupled(ibm):
    call("bl_tune_apply_coupled", ibm.handle)


def save(path=DEFAULT_PATH):
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    call("bl_tune_cache_save", str(path).encode())
    return path


def load(path=DEFAULT_PATH):
    """Merges a saved table into the process's; False if there is no file."""
    if not os.path.exists(path):
        return False
    call("bl_tune_cache_load", str(path).encode())
    return True
