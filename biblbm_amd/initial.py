"""Initial-condition pipeline over the C-ABI: the reference's
ParticlePositionInitializer (src/fsi/ParticlePositionInitializer.hh:18-86)
and create_rbc_params (progs/common/calibrate_rbc_shape.h:75-112); the RBC
volume calibration is ImmersedBoundary.shrink_rbc_volume (on the device)."""
import ctypes

import numpy as np

from . import capi
from .capi import call, ptr


def srand(seed):
    """Seeds the C library's rand(), which the position initializer draws
    from, as std::srand does in a prog."""
    ctypes.CDLL(None).srand(ctypes.c_uint(int(seed)))


class ParticlePositionInitializer:
    def __init__(self, box, boundary, max_radius):
        """box = (x0, x1, y0, y1, z0, z1) in lattice cells (a Box3D);
        boundary a walls.Boundary or None."""
        b = (ctypes.c_int * 6)(*[int(v) for v in box])
        h = ctypes.c_void_p()
        call("bl_positions_create", b, None if boundary is None else ctypes.byref(boundary), float(max_radius),
             ctypes.byref(h))
        self.handle = h

    def generate_points(self, n, radius, tag):
        """True when all n points found room (the reference's return value)."""
        try:
            call("bl_positions_generate", self.handle, int(n), float(radius), tag.encode()[:1])
            return True
        except capi.BlError:
            return False

    def get_points(self, tag):
        n = ctypes.c_longlong()
        call("bl_positions_get", self.handle, tag.encode()[:1], None, 0, ctypes.byref(n))
        out = np.zeros((n.value, 3))
        call("bl_positions_get", self.handle, tag.encode()[:1], ptr(out), n.value, ctypes.byref(n))
        return out

    def close(self):
        if getattr(self, "handle", None):
            capi.lib.bl_positions_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def create_rbc_params(rest_pos, tri, shear_modulus, K_area, K_bend):
    """create_rbc_params in the library's parameter order [shear_modulus,
    L0, k_area_global, k_area_local, k_volume, vol_desired, k_bend, theta0]:
    L0 the mean length of the mesh's links (its distinct edges, in (i0, i1)
    order), vol_desired the rest volume, the area stiffness beyond
    K() = 2 G (src/fsi/RBCParticle.hh:76-79) split 0.8 / 0.2 between global
    and local, k_volume = 1000 G / radius (largest vertex distance from the
    origin), theta0 of a regular triangulation with the mesh's vertex
    count."""
    rest_pos = np.asarray(rest_pos, float)
    tri = np.asarray(tri)
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    e = np.unique(np.sort(e, axis=1), axis=0)
    lsum = 0.0
    for i0, i1 in e:
        lsum += float(np.linalg.norm(rest_pos[i1] - rest_pos[i0]))
    a, b, c = rest_pos[tri[:, 0]], rest_pos[tri[:, 1]], rest_pos[tri[:, 2]]
    vol = np.sum(np.einsum("ij,ij->i", a, np.cross(b, c))) / 6.0
    delta = K_area - 2.0 * shear_modulus
    radius = np.sqrt((rest_pos ** 2).sum(1)).max()
    nv = len(rest_pos)
    theta0 = np.arccos((np.sqrt(3.0) * (nv - 2.0) - 5 * np.pi) / (np.sqrt(3.0) * (nv - 2.0) - 3.0 * np.pi))
    return np.array([shear_modulus, lsum / len(e), delta * 0.8, delta * 0.2, 1000.0 * shear_modulus / radius, vol,
                     K_bend, theta0])
