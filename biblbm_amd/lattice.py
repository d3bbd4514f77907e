"""Host-side lattice object over the C-ABI (bl_lattice_*).

Mirrors the parts of Palabos's MultiBlockLattice3D interface the reference
progs use for this path (src/multiBlock/multiBlockLattice3D.hh:307-347,
src/dataProcessors/dataInitializerWrapper3D.h, src/core/cell.hh:124-132):
construction with a periodicity, GuoExternalForceVOFDynamics(omega,
omega_in), bounce-back / velocity-wall cells, body force, equilibrium
initialisation, collideAndStream, computeDensity / computeVelocity.
"""
import ctypes

import numpy as np

from . import capi
from .capi import call, f64, ptr, u8

FLUID, BOUNCE_BACK, NO_DYNAMICS, VELOCITY_WALL = 0, 1, 2, 3


class Lattice:
    def __init__(self, nx, ny, nz, periodic=(True, True, True), storage="f64",
                 omega=1.0, omega_in=None):
        self.shape = (int(nx), int(ny), int(nz))
        self.ncell = nx * ny * nz
        self.storage = storage
        self.periodic = tuple(bool(p) for p in periodic)
        h = ctypes.c_void_p()
        per = np.array([int(p) for p in self.periodic], np.int32)
        call("bl_lattice_create", nx, ny, nz, ptr(per), capi.STORE[storage], ctypes.byref(h))
        self.handle = h
        self.set_dynamics(omega, omega if omega_in is None else omega_in)

    # --- configuration ---------------------------------------------------
    def set_dynamics(self, omega_out, omega_in):
        self.omega_out, self.omega_in = float(omega_out), float(omega_in)
        call("bl_lattice_set_dynamics", self.handle, self.omega_out, self.omega_in)

    def set_mask(self, mask):
        m = None if mask is None else u8(mask).reshape(-1)
        if m is not None and m.size != self.ncell:
            raise ValueError("mask must have nx*ny*nz entries")
        call("bl_lattice_set_mask", self.handle, ptr(m))

    def set_skip_inert(self, on):
        """Leave inert bounce-back cells (every neighbour bounce-back) alone
        in the collide-and-stream (bl_lattice_set_skip_inert); returns the
        number of inert cells."""
        call("bl_lattice_set_skip_inert", self.handle, int(bool(on)))
        n = ctypes.c_longlong()
        call("bl_lattice_inert_cells", self.handle, ctypes.byref(n))
        return n.value

    def set_velocity_walls(self, axis, u_lo, u_hi):
        call("bl_lattice_set_velocity_walls", self.handle, int(axis), ptr(f64(u_lo)), ptr(f64(u_hi)))

    def set_body_force(self, g):
        call("bl_lattice_set_body_force", self.handle, ptr(f64(g)))

    def set_force_field(self, F):
        F = f64(F).reshape(-1)
        if F.size != 3 * self.ncell:
            raise ValueError("force field must have nx*ny*nz*3 entries")
        call("bl_lattice_set_force_field", self.handle, ptr(F))

    def set_phase(self, phase):
        p = None if phase is None else u8(phase).reshape(-1)
        call("bl_lattice_set_phase", self.handle, ptr(p))

    def set_launch(self, block=256, xcd_remap=True):
        call("bl_lattice_set_launch", self.handle, int(block), int(bool(xcd_remap)))

    def set_launch_parity(self, parity, block=256, xcd_remap=True, nontemporal=0):
        """nontemporal: 0 / False none, 1 / True every population, 2 the
        odd step's z-aligned planes only.  A negative block -B: the odd
        step over z-segments of B cells, every access z-aligned (z
        periodic, nz a multiple of B; elsewhere, and on even steps, B)."""
        call("bl_lattice_set_launch_parity", self.handle, int(parity), int(block), int(xcd_remap), int(nontemporal))

    def launch(self, parity):
        out = np.zeros(3, np.int32)
        call("bl_lattice_get_launch", self.handle, int(parity), ptr(out))
        return dict(block=int(out[0]), xcd_remap=int(out[1]), nontemporal=int(out[2]))

    # --- state -------------------------------------------------------------
    def init_equilibrium(self, rho=None, u=None):
        r = None if rho is None else f64(rho).reshape(-1)
        v = None if u is None else f64(u).reshape(-1)
        call("bl_lattice_init_equilibrium", self.handle, ptr(r), ptr(v))

    def put_populations(self, f):
        f = f64(f).reshape(-1)
        if f.size != 19 * self.ncell:
            raise ValueError("populations must have nx*ny*nz*19 entries")
        call("bl_lattice_put_populations", self.handle, ptr(f))

    def get_populations(self):
        out = np.empty(self.shape + (19,))
        call("bl_lattice_get_populations", self.handle, ptr(out))
        return out

    def initialize(self):
        call("bl_lattice_initialize", self.handle)

    def collide_and_stream(self, n=1):
        call("bl_collide_stream", self.handle, int(n))

    def moments(self):
        rho = np.empty(self.shape)
        u = np.empty(self.shape + (3,))
        call("bl_lattice_get_moments", self.handle, ptr(rho), ptr(u))
        return rho, u

    def force(self):
        out = np.empty(self.shape + (3,))
        call("bl_lattice_get_force", self.handle, ptr(out))
        return out

    def conserved(self):
        out = np.zeros(4)
        call("bl_lattice_conserved", self.handle, ptr(out))
        return out

    @property
    def steps(self):
        s = ctypes.c_longlong()
        call("bl_lattice_steps", self.handle, ctypes.byref(s))
        return s.value

    def time_collide(self, n):
        ms = ctypes.c_double()
        call("bl_lattice_time_collide", self.handle, int(n), ctypes.byref(ms))
        return ms.value

    def save_checkpoint(self, path):
        """Populations as stored, step count and force field (bl_lattice_save_checkpoint)."""
        call("bl_lattice_save_checkpoint", self.handle, str(path).encode())

    def load_checkpoint(self, path):
        call("bl_lattice_load_checkpoint", self.handle, str(path).encode())

    def close(self):
        if getattr(self, "handle", None):
            capi.lib.bl_lattice_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
