"""Host-side immersed-boundary object over the C-ABI (bl_ibm_*, bl_shape_*,
bl_particle_*).

Mirrors the reference's ImmersedBoundaryDynamics3D
(src/fsi/ImmersedBoundaryDynamics.h:57-114): shapes live in a library
(ParticleShapeLibrary, src/fsi/ParticleShapeFactory.hh), particles are added
before init(), and one coupled step runs interpolate -> compute forces ->
spread -> move (+ voxelize).
"""
import ctypes

import numpy as np

from . import capi
from .capi import call, f64, i32, ptr

CAPSULE_SKALAK, CAPSULE_NEOHOOKEAN, RBC, SKALAK_RBC = 0, 1, 2, 3


class ImmersedBoundary:
    def __init__(self, lattice, kernel="roma"):
        self.lattice = lattice
        h = ctypes.c_void_p()
        call("bl_ibm_create", lattice.handle, capi.KERNEL[kernel], ctypes.byref(h))
        self.handle = h
        self.kernel = kernel
        self.shapes = []      # (nv, nt)
        self.particles = []   # (shape_id, nv)
        self.meshes = []      # (rest_pos, tri) per shape, as added
        self.specs = []       # (shape_id, model, params, bug_compatible) per added particle

    def add_shape(self, rest_pos, tri):
        rest_pos = f64(rest_pos).reshape(-1, 3)
        tri = i32(tri).reshape(-1, 3)
        sid = ctypes.c_int()
        call("bl_shape_add", self.handle, len(rest_pos), len(tri), ptr(rest_pos), ptr(tri), ctypes.byref(sid))
        self.shapes.append((len(rest_pos), len(tri)))
        self.meshes.append((rest_pos.copy(), tri.copy()))
        return sid.value

    def shape_info(self, sid):
        out = np.zeros(6)
        call("bl_shape_info", self.handle, int(sid), ptr(out))
        return dict(area=out[0], volume=out[1], center=out[2:5].copy(), radius=out[5])

    def add_particle(self, sid, model, params, pos, bug_compatible=True):
        params = f64(params)
        pos = f64(pos).reshape(-1, 3)
        if len(pos) != self.shapes[sid][0]:
            raise ValueError("position array does not match the shape's vertex count")
        pid = ctypes.c_int()
        call("bl_particle_add", self.handle, int(sid), int(model), ptr(params), len(params), ptr(pos),
             int(bool(bug_compatible)), ctypes.byref(pid))
        self.particles.append((sid, len(pos)))
        self.specs.append((int(sid), int(model), params.copy(), bool(bug_compatible)))
        return pid.value

    def init(self):
        call("bl_ibm_init", self.handle)

    # --- rigid particles (growth packing) ------------------------------------
    def add_rigid(self, sid, com, quat=(1.0, 0.0, 0.0, 0.0), scale=1.0, density=1.0, damping=0.0, relaxation=0.45):
        """RigidParticle3D (src/fsi/RigidParticle.hh:13-39) at centre `com`
        with orientation quaternion (w, x, y, z) and the given scale."""
        st = f64(list(com) + list(quat) + [scale, density, damping, relaxation])
        pid = ctypes.c_int()
        call("bl_rigid_add", self.handle, int(sid), ptr(st), ctypes.byref(pid))
        nv = self.shapes[sid][0]
        self.particles.append((sid, nv))
        self.specs.append((int(sid), 4, st.copy(), False))
        return pid.value

    def rigid_move(self):
        """move_vertices() of every rigid particle under the current forces."""
        call("bl_rigid_move", self.handle)

    def rigid_set_scale(self, scale, pid=-1):
        call("bl_rigid_set_scale", self.handle, int(pid), float(scale))

    def rigid_state(self, pid):
        """dict of com, vel, angvel, quat, scale, density, damping, relaxation."""
        out = np.empty(17)
        call("bl_rigid_get", self.handle, int(pid), ptr(out))
        return dict(com=out[0:3], vel=out[3:6], angvel=out[6:9], quat=out[9:13], scale=out[13], density=out[14],
                    damping=out[15], relaxation=out[16])

    def rigid_grow(self, n_iter, initial_scale, potential, boundary=None, every=10):
        """The growth loop of progs/common/create_particles.h:129-155."""
        call("bl_rigid_grow", self.handle, int(n_iter), float(initial_scale), int(every), ctypes.byref(potential),
             None if boundary is None else ctypes.byref(boundary))

    def set_boundary(self, boundary):
        """set_boundary (src/fsi/ImmersedBoundaryDynamics.h:93): near-wall
        kernel blending and vertex reflection; None removes it."""
        call("bl_ibm_set_boundary", self.handle, None if boundary is None else ctypes.byref(boundary))

    # --- interaction forces -------------------------------------------------
    def add_wall_interaction(self, boundary, potential):
        """add_force_decorator(new WallInteraction(boundary, potential))
        (src/fsi/ImmersedBoundaryDynamics.h:78, src/fsi/ForceDecorator.hh:27-35)."""
        call("bl_ibm_add_wall_interaction", self.handle, ctypes.byref(boundary), ctypes.byref(potential))

    add_force_decorator = add_wall_interaction

    def add_pp_force(self, potential):
        """add_pp_force(new PPPotentialForce(potential))
        (src/fsi/ImmersedBoundaryDynamics.h:79)."""
        call("bl_ibm_add_pp_force", self.handle, ctypes.byref(potential))

    def compute_collision_forces(self, potential, boundary=None):
        """compute_collision_forces(potential, boundary)
        (src/fsi/ImmersedBoundaryDynamics_forces.hh:505-521): vertex forces
        := collision-handler forces alone."""
        call("bl_ibm_compute_collision_forces", self.handle, ctypes.byref(potential),
             None if boundary is None else ctypes.byref(boundary))

    # --- stages ------------------------------------------------------------
    def interpolate(self):
        call("bl_ibm_interpolate", self.handle)

    def compute_forces(self):
        call("bl_ibm_compute_forces", self.handle)

    def spread(self):
        call("bl_ibm_spread", self.handle)

    def move(self):
        call("bl_ibm_move", self.handle)

    def step(self):
        call("bl_ibm_step", self.handle)

    def coupled_steps(self, n, use_graph=True):
        call("bl_coupled_steps", self.handle, int(n), int(bool(use_graph)))

    def prepare(self, n):
        """Capture the graphs coupled_steps(n) replays from the current
        state without running a step (bl_coupled_prepare)."""
        call("bl_coupled_prepare", self.handle, int(n))

    def tune(self, n_steps=8):
        """bl_tune_coupled: time each launch structure over real steps,
        keep the fastest (results do not depend on it)."""
        import json
        buf = ctypes.create_string_buffer(4096)
        call("bl_tune_coupled", self.handle, int(n_steps), buf, len(buf))
        return json.loads(buf.value.decode())

    # --- data ----------------------------------------------------------------
    def _nv(self, pid):
        return sum(n for _, n in self.particles) if pid < 0 else self.particles[pid][1]

    def _get(self, pid, what):
        out = np.empty((self._nv(pid), 3))
        call("bl_particle_get", self.handle, int(pid), int(what), ptr(out))
        return out

    def positions(self, pid=-1):
        return self._get(pid, 0)

    def velocities(self, pid=-1):
        return self._get(pid, 1)

    def forces(self, pid=-1):
        return self._get(pid, 2)

    def set_positions(self, pos, pid=-1):
        pos = f64(pos).reshape(-1, 3)
        if len(pos) != self._nv(pid):
            raise ValueError("position array size mismatch")
        call("bl_particle_set_positions", self.handle, int(pid), ptr(pos))

    def observables(self, pid):
        out = np.zeros(14)
        call("bl_particle_observables", self.handle, int(pid), ptr(out))
        return dict(area=out[0], volume=out[1], com=out[2:5].copy(), inertia=out[5:14].reshape(3, 3).copy())

    def phase(self):
        out = np.empty(self.lattice.shape, np.uint8)
        call("bl_ibm_get_phase", self.handle, ptr(out))
        return out

    # --- initial conditions ----------------------------------------------------
    def shrink_rbc_volume(self, vol_final, n_it, damping=1e-2, pid=-1):
        """shrink_rbc_volume (progs/common/calibrate_rbc_shape.h:13-72) on the
        device for particle pid or every volume-constrained particle; returns
        the damping each particle finished with."""
        out = np.zeros(max(len(self.particles), 1))
        call("bl_rbc_shrink_volume", self.handle, int(pid), float(vol_final), int(n_it), float(damping), ptr(out))
        return out[:len(self.particles)]

    # --- on-disk formats ------------------------------------------------------
    def set_shape_tag(self, sid, tag):
        call("bl_shape_set_tag", self.handle, int(sid), tag.encode())

    def describe(self, pid):
        """(shape id, membrane model, the reference's particle id)."""
        sid, model, pid_ref = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
        call("bl_particle_describe", self.handle, int(pid), ctypes.byref(sid), ctypes.byref(model),
             ctypes.byref(pid_ref))
        return sid.value, model.value, pid_ref.value

    def pack(self, pid):
        """ParticleBase3D::pack bytes of one particle (bl_particle_pack)."""
        n = ctypes.c_longlong()
        call("bl_particle_pack", self.handle, int(pid), None, 0, ctypes.byref(n))
        buf = ctypes.create_string_buffer(n.value)
        call("bl_particle_pack", self.handle, int(pid), buf, n.value, ctypes.byref(n))
        return buf.raw[:n.value]

    def save_checkpoint(self, folder, name="checkpoint_fsi"):
        """save_checkpoint (src/fsi/ImmersedBoundaryDynamics_io.hh:17-73);
        creates <folder>/<name>/ for the rank file."""
        import os
        os.makedirs(os.path.join(folder, name), exist_ok=True)
        call("bl_ibm_save_checkpoint", self.handle, folder.encode(), name.encode())

    def load_checkpoint(self, folder, name="checkpoint_fsi"):
        """load_checkpoint (src/fsi/ImmersedBoundaryDynamics_io.hh:75-158):
        replaces the particles; call before init()."""
        call("bl_ibm_load_checkpoint", self.handle, folder.encode(), name.encode())
        n = ctypes.c_int()
        call("bl_ibm_count", self.handle, ctypes.byref(n), None)
        self.particles = []
        for p in range(n.value):
            sid = self.describe(p)[0]
            self.particles.append((sid, self.shapes[sid][0]))

    def write_vtk(self, pid, path):
        call("bl_particle_write_vtk", self.handle, int(pid), str(path).encode())

    def write_gmsh(self, pid, path):
        call("bl_particle_write_gmsh", self.handle, int(pid), str(path).encode())

    def write_lightweight(self, path, iteration):
        call("bl_ibm_write_lightweight", self.handle, str(path).encode(), int(iteration))

    def set_fused_staging(self, mode):
        """-1 auto, 0 a staging pass of its own, 1 staging carried by the
        collide-and-stream launch (bl_ibm_set_fused_staging)."""
        call("bl_ibm_set_fused_staging", self.handle, int(mode))

    def set_layered(self, mode):
        """-1 auto, 0 membrane forces and spreading on a second stream, 1
        one stream with paired stages (bl_ibm_set_layered)."""
        call("bl_ibm_set_layered", self.handle, int(mode))

    LAUNCH_KEYS = ("bin_block", "elem_block", "gather_lanes", "interp_block", "spread_small", "spread_big", "xcd")

    def set_launch(self, **params):
        """Immersed-boundary launch parameters (bl_ibm_set_launch), by name;
        the ones not given keep their values."""
        cur = self.launch()
        for k, v in params.items():
            if k not in cur:
                raise KeyError(k)
            cur[k] = int(v)
        arr = np.array([cur[k] for k in self.LAUNCH_KEYS], np.int32)
        call("bl_ibm_set_launch", self.handle, ptr(arr))

    def launch(self):
        out = np.zeros(len(self.LAUNCH_KEYS), np.int32)
        call("bl_ibm_get_launch", self.handle, ptr(out))
        return dict(zip(self.LAUNCH_KEYS, (int(x) for x in out)))

    def tune_launch(self, n_steps=8):
        """bl_tune_ibm: coordinate descent over the launch parameters with
        real steps; keeps and records the fastest."""
        import json
        buf = ctypes.create_string_buffer(16384)
        call("bl_tune_ibm", self.handle, int(n_steps), buf, len(buf))
        return json.loads(buf.value.decode())

    def tile_stats(self):
        out = np.zeros(2, np.int64)
        call("bl_ibm_tile_stats", self.handle, ptr(out))
        return dict(active_tiles=int(out[0]), list_entries=int(out[1]))

    def stage_times(self):
        out = np.zeros(5)
        call("bl_ibm_stage_times", self.handle, ptr(out))
        return dict(zip(["bin", "compute_forces", "interpolate_spread", "move", "voxelize"], out))

    def close(self):
        if getattr(self, "handle", None):
            capi.lib.bl_ibm_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def taylor_deformation(inertia):
    """Taylor deformation index D = (L - B) / (L + B) of the ellipsoid with
    the same inertia tensor (principal moments I1 <= I2 <= I3 give semi-axes
    a_k^2 proportional to the sum of the other two moments minus itself)."""
    w = np.linalg.eigvalsh(0.5 * (inertia + inertia.T))
    s = w.sum()
    axes2 = np.clip(s / 2.0 - w, 0, None)   # proportional to a_k^2
    a = np.sqrt(axes2)
    L, B = a.max(), a.min()
    return (L - B) / (L + B)
