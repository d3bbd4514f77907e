"""Interaction potentials over the C-ABI (bl_potential).

Mirrors src/fsi/Potentials.h: MorsePotential(r_min, beta, strength)
(:28-56) and SpringPotential(min_dist, strength) (:58-72), each callable
for its force at a distance and carrying its cut-off.  They are handed to
ImmersedBoundary.add_wall_interaction / add_pp_force /
compute_collision_forces, which evaluate them on the GPU.
"""
import ctypes

from .capi import call

MORSE, SPRING = 0, 1


class Potential(ctypes.Structure):
    """bl_potential: kind MORSE (r_min, beta, strength) or SPRING
    (min_dist, strength)."""
    _fields_ = [("kind", ctypes.c_int), ("p", ctypes.c_double * 3)]

    def __call__(self, dist):
        f = ctypes.c_double()
        call("bl_potential_eval", ctypes.byref(self), float(dist), ctypes.byref(f), None)
        return f.value

    def get_cutoff_distance(self):
        c = ctypes.c_double()
        call("bl_potential_eval", ctypes.byref(self), 0.0, None, ctypes.byref(c))
        return c.value

    get_repulsion_distance = get_cutoff_distance


def MorsePotential(r_min, beta, strength):
    """Force 2 beta strength x (x - 1), x = exp(beta (r_min - max(d, 0))),
    cut-off r_min."""
    return Potential(MORSE, (ctypes.c_double * 3)(r_min, beta, strength))


def SpringPotential(min_dist, strength):
    """Force (min_dist - d) strength, cut-off min_dist."""
    return Potential(SPRING, (ctypes.c_double * 3)(min_dist, strength, 0.0))
