"""Ensembles of independent boxes over the C-ABI (bl_ensemble_*), and the
collide launch-shape tuner (bl_tune_collide).

No counterpart in the reference, which runs each box (one rbc_in_shear
setup) as its own program: here every box keeps its own Lattice /
ImmersedBoundary objects and one hipGraph launch advances all of them.
"""
import ctypes
import json

from . import capi
from .capi import call


class Ensemble:
    def __init__(self, boxes):
        """boxes: list of (Lattice, ImmersedBoundary or None); the objects
        must stay alive as long as the ensemble."""
        self.boxes = list(boxes)
        n = len(self.boxes)
        lats = (ctypes.c_void_p * n)(*[b[0].handle.value for b in self.boxes])
        ibms = (ctypes.c_void_p * n)(*[(b[1].handle.value if b[1] is not None else None) for b in self.boxes])
        h = ctypes.c_void_p()
        call("bl_ensemble_create", n, ctypes.cast(lats, ctypes.c_void_p), ctypes.cast(ibms, ctypes.c_void_p), 0,
             ctypes.byref(h))
        self.handle = h

    def steps(self, n, use_graph=True):
        call("bl_ensemble_steps", self.handle, int(n), int(bool(use_graph)))

    def set_batched(self, mode):
        """-1 batched when the boxes allow it (default), 0 one graph branch
        per box, 1 batched only (bl_ensemble_set_batched)."""
        call("bl_ensemble_set_batched", self.handle, int(mode))

    @property
    def batched(self):
        out = ctypes.c_int()
        call("bl_ensemble_is_batched", self.handle, ctypes.byref(out))
        return bool(out.value)

    def tune(self, n_steps=4):
        """bl_ensemble_tune: the batched collide's workgroup order and
        nontemporal level per AA parity, over real steps; returns the
        report."""
        buf = ctypes.create_string_buffer(8192)
        call("bl_ensemble_tune", self.handle, int(n_steps), buf, len(buf))
        return json.loads(buf.value.decode())

    def set_launch(self, parity, order, nontemporal=0):
        """bl_ensemble_set_launch: the batched collide's workgroup order and
        nontemporal level for steps of one AA parity, without tuning."""
        call("bl_ensemble_set_launch", self.handle, int(parity), int(order), int(nontemporal))

    def time_steps(self, n, use_graph=True):
        ms = ctypes.c_double()
        call("bl_ensemble_time_steps", self.handle, int(n), int(bool(use_graph)), ctypes.byref(ms))
        return ms.value

    def close(self):
        if getattr(self, "handle", None):
            capi.lib.bl_ensemble_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tune_collide(lattice):
    """Times every launch shape of the lattice's collide kernel per step
    parity, keeps the fastest (the lattice state is left as it was);
    returns the report (biblbm_amd.tuning.tune_collide)."""
    from .tuning import tune_collide as _t
    return _t(lattice)
