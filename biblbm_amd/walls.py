"""Curved walls over the C-ABI: analytic boundaries, the Guo curved wall and
the boundary mask helper.

Mirrors the fsi classes the pipe and channel progs use:
PipeBoundary / ParallelPlatesBoundary (src/fsi/Boundary.h:132-365),
GuoRigidWallBoundary (src/fsi/GuoBoundaryCondition.h:19-52) and
defineDynamics(..., BoundaryOutside(boundary, margin), NoDynamics)
(src/fsi/Boundary.h:114-128, progs/rbc_in_pipe/rbc_in_pipe.cpp:302-318).
"""
import ctypes

import numpy as np

from . import capi
from .capi import call, f64, ptr

PIPE, PLATES = 0, 1


class Boundary(ctypes.Structure):
    """bl_boundary: kind PIPE (y0, z0, radius, length) along x, or PLATES
    (y0, y1) normal to y."""
    _fields_ = [("kind", ctypes.c_int), ("p", ctypes.c_double * 4)]

    @classmethod
    def pipe(cls, y0, z0, radius, length=0.0):
        return cls(PIPE, (ctypes.c_double * 4)(y0, z0, radius, length))

    @classmethod
    def plates(cls, y0, y1):
        return cls(PLATES, (ctypes.c_double * 4)(y0, y1, 0.0, 0.0))


def outside_mask(boundary, shape, margin=1.0, mask=None, value=2):
    """Sets `value` (default: no dynamics) on every cell with
    !contains(cell, -margin).  The reference passes sqrt(2) to an integer
    margin, i.e. 1."""
    nx, ny, nz = shape
    m = np.zeros(shape, np.uint8) if mask is None else np.ascontiguousarray(mask, np.uint8).copy()
    call("bl_boundary_outside_mask", ctypes.byref(boundary), nx, ny, nz, float(margin), ptr(m), value)
    return m


class GuoWall:
    """GuoRigidWallBoundary(boundary, lattice): wall nodes found at creation;
    insert() makes the extrapolation run after every collide-and-stream."""

    def __init__(self, lattice, boundary):
        self.lattice = lattice
        self.boundary = boundary
        h = ctypes.c_void_p()
        call("bl_guo_wall_create", lattice.handle, ctypes.byref(boundary), ctypes.byref(h))
        self.handle = h
        n = ctypes.c_longlong()
        call("bl_guo_wall_count", h, ctypes.byref(n))
        self.n = n.value

    def nodes(self):
        cell = np.empty(self.n, np.int64)
        d = np.empty(self.n, np.int32)
        delta = np.empty(self.n)
        wp = np.empty((self.n, 3))
        call("bl_guo_wall_nodes", self.handle, ptr(cell), ptr(d), ptr(delta), ptr(wp))
        return dict(cell=cell, dir=d, delta=delta, wall_pos=wp)

    def set_wall_velocity(self, vel):
        """One velocity for every node, or a function of the wall position."""
        if callable(vel):
            wp = self.nodes()["wall_pos"]
            v = f64(np.array([vel(p) for p in wp], float).reshape(-1, 3))
            call("bl_guo_wall_set_node_velocities", self.handle, ptr(v))
        else:
            v = f64(np.asarray(vel, float).reshape(3))
            call("bl_guo_wall_set_velocity", self.handle, ptr(v))

    def set_node_velocities(self, vel):
        v = f64(np.asarray(vel, float).reshape(self.n, 3))
        call("bl_guo_wall_set_node_velocities", self.handle, ptr(v))

    def insert(self):
        call("bl_guo_wall_insert", self.handle)

    def close(self):
        if self.handle:
            call("bl_guo_wall_destroy", self.handle)
            self.handle = None
