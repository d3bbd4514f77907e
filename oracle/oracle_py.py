"""ctypes binding of the FP64 CPU restatement in oracle/liboracle.so.

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product package.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int
_L = ctypes.c_long

_SIGS = {
    "or_collide_cell": (None, [_P, _P, _D, _D, _D]),
    "or_lbm_step": (None, [_I, _I, _I, _P, _D, _D, _P, _P, _P, _P, _P, _I, _P]),
    "or_velocity_walls": (None, [_I, _I, _I, _P, _D, _P, _P, _P, _I, _P]),
    "or_lbm_moments": (None, [_I, _I, _I, _P, _P, _P, _I, _P, _P, _P]),
    "or_lbm_init_equilibrium": (None, [_L, _P, _P, _P]),
    "or_lbm_conserved": (None, [_L, _P, _P, _P]),
    "or_dirac_1d": (_D, [_I, _D]),
    "or_dirac_support": (_I, [_I]),
    "or_dirac_half_support": (_D, [_I]),
    "or_support_base": (_L, [_I, _D]),
    "or_interpolate": (None, [_I, _I, _I, _I, _P, _P, _L, _P, _P]),
    "or_spread": (None, [_I, _I, _I, _I, _P, _P, _L, _P, _P]),
    "or_capsule_forces": (None, [_L, _L, _P, _P, _P, _D, _D, _I, _I, _P]),
    "or_build_links": (_L, [_L, _L, _P, _P, _P, _P, _P]),
    "or_build_pairs": (_L, [_L, _L, _P, _P, _L, _P, _P, _P, _P]),
    "or_rbc_forces": (None, [_L, _L, _P, _P, _L, _P, _P, _L, _P, _P, _D, _P, _D, _D, _P, _P]),
    "or_particle_update": (None, [_L, _L, _P, _P, _P, _P]),
    "or_moment_of_inertia": (None, [_L, _L, _P, _P, _P]),
    "or_ray_hits_z": (_L, [_L, _L, _P, _P, _L, _P]),
    "or_voxel_fill": (None, [_I, _I, _I, _P, _L, _P, _P]),
    "or_boundary_contains": (_I, [_P, _P, _D]),
    "or_boundary_less_than": (_I, [_P, _P, _D]),
    "or_boundary_distance": (_D, [_P, _P]),
    "or_boundary_trace_ray": (_I, [_P, _P, _P, _P]),
    "or_boundary_normal": (None, [_P, _P, _P]),
    "or_guo_wall_nodes": (_L, [_I, _I, _I, _P, _L, _P, _P, _P, _P]),
    "or_guo_wall_apply": (None, [_I, _I, _I, _P, _P, _P, _L, _P, _P, _P, _P]),
    "or_interpolate_near": (None, [_I, _I, _I, _I, _P, _P, _L, _P, _P, _P]),
    "or_spread_near": (None, [_I, _I, _I, _I, _P, _P, _L, _P, _P, _P]),
    "or_move_reflect": (None, [_L, _P, _P, _P]),
    "or_potential_force": (_D, [_P, _D]),
    "or_potential_cutoff": (_D, [_P]),
    "or_wall_interaction": (None, [_P, _P, _L, _P, _P]),
    "or_collision_forces": (_L, [_I, _I, _I, _P, _P, _L, _P, _P, _P, _P]),
    "or_boundary_intersects_rect": (_I, [_P, _P, _D]),
}


class RbcParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "shear_modulus", "L0", "k_area_global", "k_area_local",
        "k_volume", "vol_desired", "k_bend", "theta0")]


def _load():
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _p(a):
    return a.ctypes.data_as(_P)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


# ----------------------------------------------------------------- lattice
def lbm_run(f, force, phase, mask, periodic, omega_out, omega_in, steps,
            wall_axis=1, wall_u=None, guo=None):
    """Advance AoS populations f (nx,ny,nz,19) by `steps` collide-and-stream
    steps; returns a new array.  `guo`: (wall nodes, wall velocities) of a
    Guo curved wall applied after every step."""
    if guo is not None:
        for _ in range(steps):
            f = lbm_run(f, force, phase, mask, periodic, omega_out, omega_in, 1, wall_axis, wall_u)
            f = guo_wall_apply(f, periodic, guo[0], guo[1], mask)
        return f
    f = _c(f, np.float64).copy()
    nx, ny, nz = f.shape[:3]
    tmp = np.empty_like(f)
    force = _c(force, np.float64)
    phase = _c(phase, np.float64)
    mask = _c(mask, np.uint8)
    per = _c(periodic, np.int32)
    wu = _c(np.zeros(6) if wall_u is None else wall_u, np.float64)
    for _ in range(steps):
        lib.or_lbm_step(nx, ny, nz, _p(per), omega_out, omega_in, _p(f), _p(tmp),
                        _p(force), _p(phase), _p(mask), wall_axis, _p(wu))
    return f


class Boundary(ctypes.Structure):
    """Analytic boundary: kind 0 pipe along x (y0, z0, radius, length),
    kind 1 parallel plates y0 < y < y1."""
    _fields_ = [("kind", ctypes.c_int), ("p", ctypes.c_double * 4)]

    @classmethod
    def pipe(cls, y0, z0, radius, length=0.0):
        return cls(0, (ctypes.c_double * 4)(y0, z0, radius, length))

    @classmethod
    def plates(cls, y0, y1):
        return cls(1, (ctypes.c_double * 4)(y0, y1, 0.0, 0.0))


def guo_wall_nodes(shape, boundary):
    """Wall nodes of the Guo curved wall: dict(cell, dir, delta, wall_pos)."""
    nx, ny, nz = shape
    n = lib.or_guo_wall_nodes(nx, ny, nz, ctypes.byref(boundary), 0, None, None, None, None)
    cell = np.empty(n, np.int64)
    d = np.empty(n, np.int32)
    delta = np.empty(n)
    wp = np.empty((n, 3))
    lib.or_guo_wall_nodes(nx, ny, nz, ctypes.byref(boundary), n, _p(cell), _p(d), _p(delta), _p(wp))
    return dict(cell=cell, dir=d, delta=delta, wall_pos=wp)


def guo_wall_apply(f, periodic, walls, vel, mask=None):
    """The Guo extrapolation over the wall nodes (in place on a copy)."""
    f = _c(f, np.float64).copy()
    nx, ny, nz = f.shape[:3]
    vel = _c(np.broadcast_to(np.asarray(vel, float), (len(walls["cell"]), 3)), np.float64)
    m = None if mask is None else _p(_c(mask, np.uint8))
    lib.or_guo_wall_apply(nx, ny, nz, _p(_c(periodic, np.int32)), _p(f), m, len(walls["cell"]),
                          _p(_c(walls["cell"], np.int64)), _p(_c(walls["dir"], np.int32)),
                          _p(_c(walls["delta"], np.float64)), _p(vel))
    return f


def lbm_velocity_walls(f, force, mask, periodic, omega_out, wall_axis=1, wall_u=None):
    f = _c(f, np.float64).copy()
    nx, ny, nz = f.shape[:3]
    wu = _c(np.zeros(6) if wall_u is None else wall_u, np.float64)
    lib.or_velocity_walls(nx, ny, nz, _p(_c(periodic, np.int32)), omega_out, _p(f),
                          _p(_c(force, np.float64)), _p(_c(mask, np.uint8)), wall_axis, _p(wu))
    return f


def lbm_moments(f, force, mask=None, wall_axis=1, wall_u=None):
    f = _c(f, np.float64)
    nx, ny, nz = f.shape[:3]
    mask = np.zeros((nx, ny, nz), np.uint8) if mask is None else _c(mask, np.uint8)
    wu = _c(np.zeros(6) if wall_u is None else wall_u, np.float64)
    rho = np.empty((nx, ny, nz))
    u = np.empty((nx, ny, nz, 3))
    lib.or_lbm_moments(nx, ny, nz, _p(f), _p(_c(force, np.float64)), _p(mask), wall_axis,
                       _p(wu), _p(rho), _p(u))
    return rho, u


def lbm_init_equilibrium(rho, u):
    rho = _c(rho, np.float64)
    u = _c(u, np.float64)
    f = np.empty(rho.shape + (19,))
    lib.or_lbm_init_equilibrium(rho.size, _p(rho), _p(u), _p(f))
    return f


def lbm_conserved(f, mask=None):
    f = _c(f, np.float64)
    ncell = f.size // 19
    out = np.zeros(4)
    m = None if mask is None else _c(mask, np.uint8)
    lib.or_lbm_conserved(ncell, _p(f), None if m is None else _p(m), _p(out))
    return out


def collide_cell(f, F, phase, omega_out, omega_in):
    g = _c(f, np.float64).copy()
    lib.or_collide_cell(_p(g), _p(_c(F, np.float64)), phase, omega_out, omega_in)
    return g


# ------------------------------------------------------- immersed boundary
KERNELS = {"roma": 0, "peskin": 1, "tophat": 2}


def dirac(kernel, r):
    k = KERNELS.get(kernel, kernel)
    return np.array([lib.or_dirac_1d(k, float(v)) for v in np.ravel(r)]).reshape(np.shape(r))


def support_base(kernel, x):
    k = KERNELS.get(kernel, kernel)
    return np.array([lib.or_support_base(k, float(v)) for v in np.ravel(x)],
                    dtype=np.int64).reshape(np.shape(x))


def interpolate(kernel, ufield, periodic, pos):
    k = KERNELS.get(kernel, kernel)
    u = _c(ufield, np.float64)
    pos = _c(pos, np.float64)
    nx, ny, nz = u.shape[:3]
    vel = np.empty((len(pos), 3))
    lib.or_interpolate(k, nx, ny, nz, _p(_c(periodic, np.int32)), _p(u), len(pos), _p(pos), _p(vel))
    return vel


def spread(kernel, shape, periodic, pos, fvert, base=None):
    k = KERNELS.get(kernel, kernel)
    field = np.zeros(tuple(shape) + (3,)) if base is None else _c(base, np.float64).copy()
    pos = _c(pos, np.float64)
    lib.or_spread(k, shape[0], shape[1], shape[2], _p(_c(periodic, np.int32)), _p(field),
                  len(pos), _p(pos), _p(_c(fvert, np.float64)))
    return field


def interpolate_near(kernel, ufield, periodic, pos, boundary):
    """Interpolation with the near-wall blend (boundary None: plain)."""
    k = KERNELS.get(kernel, kernel)
    u = _c(ufield, np.float64)
    pos = _c(pos, np.float64)
    nx, ny, nz = u.shape[:3]
    vel = np.empty((len(pos), 3))
    b = None if boundary is None else ctypes.byref(boundary)
    lib.or_interpolate_near(k, nx, ny, nz, _p(_c(periodic, np.int32)), _p(u), len(pos), _p(pos), b, _p(vel))
    return vel


def spread_near(kernel, shape, periodic, pos, fvert, boundary, base=None):
    """Spreading of all vertices, bulk vertices first, then the near-wall
    ones with the blend (boundary None: plain, in vertex order)."""
    k = KERNELS.get(kernel, kernel)
    field = np.zeros(tuple(shape) + (3,)) if base is None else _c(base, np.float64).copy()
    pos = _c(pos, np.float64)
    b = None if boundary is None else ctypes.byref(boundary)
    lib.or_spread_near(k, shape[0], shape[1], shape[2], _p(_c(periodic, np.int32)), _p(field), len(pos),
                       _p(pos), _p(_c(fvert, np.float64)), b)
    return field


class Potential(ctypes.Structure):
    """Interaction potential: kind 0 Morse (r_min, beta, strength), kind 1
    spring (min_dist, strength)."""
    _fields_ = [("kind", ctypes.c_int), ("p", ctypes.c_double * 3)]

    @classmethod
    def morse(cls, r_min, beta, strength):
        return cls(0, (ctypes.c_double * 3)(r_min, beta, strength))

    @classmethod
    def spring(cls, min_dist, strength):
        return cls(1, (ctypes.c_double * 3)(min_dist, strength, 0.0))


def potential_force(u, dist):
    return lib.or_potential_force(ctypes.byref(u), float(dist))


def wall_interaction(boundary, u, pos, force=None):
    """Wall interaction decorator: returns force - F(d) n (per vertex)."""
    pos = _c(pos, np.float64)
    out = np.zeros_like(pos) if force is None else _c(force, np.float64).copy()
    lib.or_wall_interaction(ctypes.byref(boundary), ctypes.byref(u), len(pos), _p(pos), _p(out))
    return out


def collision_forces(shape, periodic, u, pos, obj, wall=None, force=None):
    """Particle-particle interaction forces (collision handler over the
    lattice box), added to `force`; returns (force, interacting pairs)."""
    pos = _c(pos, np.float64)
    out = np.zeros_like(pos) if force is None else _c(force, np.float64).copy()
    w = None if wall is None else ctypes.byref(wall)
    n = lib.or_collision_forces(shape[0], shape[1], shape[2], _p(_c(periodic, np.int32)), ctypes.byref(u),
                                len(pos), _p(pos), _p(_c(obj, np.int32)), w, _p(out))
    return out, n


def move_reflect(pos, vel, boundary):
    pos = _c(pos, np.float64).copy()
    lib.or_move_reflect(len(pos), _p(pos), _p(_c(vel, np.float64)), ctypes.byref(boundary))
    return pos


def capsule_forces(tri, rest_pos, pos, G, C, model="skalak", bug_compatible=True):
    tri = _c(tri, np.int32)
    rest_pos = _c(rest_pos, np.float64)
    pos = _c(pos, np.float64)
    F = np.zeros_like(pos)
    lib.or_capsule_forces(len(pos), len(tri), _p(tri), _p(rest_pos), _p(pos), G, C,
                          0 if model == "skalak" else 1, int(bug_compatible), _p(F))
    return F


def build_topology(tri, rest_pos):
    tri = _c(tri, np.int32)
    rest_pos = _c(rest_pos, np.float64)
    nv, nt = len(rest_pos), len(tri)
    li = np.zeros((3 * nt, 2), np.int32)
    ll = np.zeros(3 * nt)
    lt = np.zeros((3 * nt, 2), np.int32)
    nl = lib.or_build_links(nv, nt, _p(tri), _p(rest_pos), _p(li), _p(ll), _p(lt))
    pi = np.zeros((3 * nt, 4), np.int32)
    pc = np.zeros((3 * nt, 2))
    npair = lib.or_build_pairs(nv, nt, _p(tri), _p(rest_pos), nl, _p(li), _p(lt), _p(pi), _p(pc))
    return dict(link_i=li[:nl].copy(), link_len=ll[:nl].copy(), link_t=lt[:nl].copy(),
                pair_i=pi[:npair].copy(), pair_cs0=pc[:npair].copy())


def rbc_forces(tri, tri_area0, topo, area0_total, params, area, volume, pos):
    tri = _c(tri, np.int32)
    pos = _c(pos, np.float64)
    F = np.zeros_like(pos)
    p = RbcParams(*[float(v) for v in params])
    li, ll = _c(topo["link_i"], np.int32), _c(topo["link_len"], np.float64)
    pi, pc = _c(topo["pair_i"], np.int32), _c(topo["pair_cs0"], np.float64)
    lib.or_rbc_forces(len(pos), len(tri), _p(tri), _p(_c(tri_area0, np.float64)),
                      len(ll), _p(li), _p(ll), len(pc), _p(pi), _p(pc), area0_total,
                      ctypes.byref(p), area, volume, _p(pos), _p(F))
    return F


def particle_update(tri, pos):
    tri = _c(tri, np.int32)
    pos = _c(pos, np.float64)
    out = np.zeros(5)
    bb = np.zeros(6)
    lib.or_particle_update(len(pos), len(tri), _p(tri), _p(pos), _p(out), _p(bb))
    return dict(area=out[0], volume=out[1], com=out[2:5].copy(), bbox=bb)


def moment_of_inertia(tri, pos):
    tri = _c(tri, np.int32)
    pos = _c(pos, np.float64)
    out = np.zeros(9)
    lib.or_moment_of_inertia(len(pos), len(tri), _p(tri), _p(pos), _p(out))
    return out.reshape(3, 3)


def ray_hits_z(tri, pos, cap=1 << 20):
    tri = _c(tri, np.int32)
    pos = _c(pos, np.float64)
    hits = np.zeros((cap, 3))
    n = lib.or_ray_hits_z(len(pos), len(tri), _p(tri), _p(pos), cap, _p(hits))
    return hits[:min(n, cap)].copy()


def voxel_fill(shape, periodic, hits):
    hits = _c(hits, np.float64).reshape(-1, 3)
    phase = np.zeros(tuple(shape), np.uint8)
    lib.or_voxel_fill(shape[0], shape[1], shape[2], _p(_c(periodic, np.int32)), len(hits), _p(hits),
                      _p(phase))
    return phase


# ------------------------------------------------------- composed steps
class Particle:
    """Oracle-side particle: rest mesh, model, parameters, vertex state."""

    def __init__(self, rest, tri, model, params, pos, bug_compatible=True):
        self.rest = _c(rest, np.float64)
        self.tri = _c(tri, np.int32)
        self.model = model
        self.params = np.asarray(params, float)
        self.pos = _c(pos, np.float64).copy()
        self.bug = bug_compatible
        self.topo = build_topology(self.tri, self.rest)
        a, b, c = self.rest[self.tri[:, 0]], self.rest[self.tri[:, 1]], self.rest[self.tri[:, 2]]
        self.tri_area0 = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)
        self.area0 = self.tri_area0.sum()
        g = particle_update(self.tri, self.pos)
        self.area, self.volume, self.com = g["area"], g["volume"], g["com"]

    def forces(self):
        m, P = self.model, self.params
        F = np.zeros_like(self.pos)
        if m in (0, 1, 3):
            F += capsule_forces(self.tri, self.rest, self.pos, P[0], P[1],
                                "neohookean" if m == 1 else "skalak", self.bug)
        if m == 2:
            F += rbc_forces(self.tri, self.tri_area0, self.topo, self.area0, P, self.area, self.volume,
                            self.pos)
        if m == 3:
            # area / volume / bending terms of the RBC model without its links
            topo = dict(self.topo)
            topo["link_i"] = np.zeros((0, 2), np.int32)
            topo["link_len"] = np.zeros(0)
            rp = np.array([1.0, 1.0, P[2], P[3], P[4], P[5], P[6], 0.0])
            F += rbc_forces(self.tri, self.tri_area0, topo, self.area0, rp, self.area, self.volume,
                            self.pos)
        return F

    def move(self, vel, shape, periodic, boundary=None):
        if boundary is None:
            self.pos += vel
        else:
            self.pos = move_reflect(self.pos, vel, boundary)
        g = particle_update(self.tri, self.pos)
        self.area, self.volume = g["area"], g["volume"]
        com = g["com"].copy()
        for d in range(3):
            if periodic[d]:
                L = float(shape[d])
                r = com[d] - L * np.floor((com[d] + 0.5) * (1.0 / L))
                self.pos[:, d] += r - com[d]
                com[d] = r
        self.com = com


def vertex_forces(particles, shape, periodic, wall_interactions=(), pp_forces=()):
    """compute_forces of the immersed boundary: per particle the wall
    decorators (from zero) then the membrane forces, then the
    particle-particle forces over all vertices (particle index as the
    owner).  Returns one (nv, 3) array per particle."""
    forces = []
    for p in particles:
        F = p.forces()
        if wall_interactions:
            fw = np.zeros_like(p.pos)
            for b, u in wall_interactions:
                fw = wall_interaction(b, u, p.pos, fw)
            F = fw + F
        forces.append(F)
    if pp_forces and particles:
        allpos = np.concatenate([p.pos for p in particles])
        obj = np.concatenate([np.full(len(p.pos), i, np.int32) for i, p in enumerate(particles)])
        allF = np.concatenate(forces)
        for u in pp_forces:
            allF, _ = collision_forces(shape, periodic, u, allpos, obj, None, allF)
        forces = np.split(allF, np.cumsum([len(p.pos) for p in particles])[:-1])
    return forces


def coupled_step(f, force, phase, mask, periodic, omega, omega_in, body, particles, kernel,
                 wall_axis=1, wall_u=None, boundary=None, guo=None, wall_interactions=(), pp_forces=()):
    """One time step of the progs' loop, composed from the oracle's pieces:
    computeVelocity (u = j/rho + F/2 with the last collision's force), force
    reset to the body force, interpolation, vertex forces (wall decorators,
    membrane, particle-particle), spreading, vertex advection (+ COM
    remap), then collide-and-stream with the new force.  With a `boundary`
    (set_boundary), near-wall vertices blend with the top hat and are spread
    after all others, and advected vertices reflect off it; `guo` = (wall
    nodes, velocities) of a Guo curved wall run after the
    collide-and-stream.  Returns (f, force, per-particle velocities)."""
    shape = f.shape[:3]
    _, u = lbm_moments(f, force, mask, wall_axis, wall_u)
    new_force = np.broadcast_to(np.asarray(body, float), tuple(shape) + (3,)).copy()
    pforces = vertex_forces(particles, shape, periodic, wall_interactions, pp_forces)
    if boundary is None:
        vels = []
        for p in particles:
            vels.append(interpolate(kernel, u, periodic, p.pos))
        for p, F in zip(particles, pforces):
            new_force = spread(kernel, shape, periodic, p.pos, F, base=new_force)
    else:
        allpos = np.concatenate([p.pos for p in particles])
        allv = interpolate_near(kernel, u, periodic, allpos, boundary)
        forces = np.concatenate(pforces)
        new_force = spread_near(kernel, shape, periodic, allpos, forces, boundary, base=new_force)
        cuts = np.cumsum([len(p.pos) for p in particles])[:-1]
        vels = np.split(allv, cuts)
    for p, v in zip(particles, vels):
        p.move(v, shape, periodic, boundary)
    f = lbm_run(f, new_force, phase, mask, periodic, omega, omega_in, 1, wall_axis, wall_u, guo=guo)
    return f, new_force, vels


# ------------------------------------------------- initial conditions
def create_rbc_params(rest, tri, shear_modulus, K_area, K_bend):
    """create_rbc_params (progs/common/calibrate_rbc_shape.h:75-112), in the
    library's parameter order [shear_modulus, L0, k_area_global,
    k_area_local, k_volume, vol_desired, k_bend, theta0]: L0 the mean link
    length, vol_desired the rest volume, the area stiffness beyond
    K() = 2 G (src/fsi/RBCParticle.hh:76-79) split 0.8 / 0.2 between global
    and local, k_volume = 1000 G / radius (largest vertex distance from the
    origin), theta0 of a regular icosahedral triangulation."""
    rest = np.asarray(rest, float)
    topo = build_topology(tri, rest)
    lavg = topo["link_len"].sum() / len(topo["link_len"])
    vol = particle_update(tri, rest)["volume"]
    delta = K_area - 2.0 * shear_modulus
    radius = np.sqrt((rest ** 2).sum(1)).max()
    nv = len(rest)
    theta0 = np.arccos((np.sqrt(3.0) * (nv - 2.0) - 5 * np.pi) / (np.sqrt(3.0) * (nv - 2.0) - 3.0 * np.pi))
    return np.array([shear_modulus, lavg, delta * 0.8, delta * 0.2, 1000.0 * shear_modulus / radius, vol, K_bend,
                     theta0])


def node_masses(tri, pos):
    """Area / 3 of every incident triangle, summed per vertex in triangle
    order (RBCParticle::relax_nodes, src/fsi/RBCParticle.hh:322-330;
    tri::triangle_area, src/fsi/TriangleUtils.h:52-55)."""
    M = np.zeros(len(pos))
    for t in np.asarray(tri):
        a = 0.5 * np.linalg.norm(np.cross(pos[t[1]] - pos[t[0]], pos[t[2]] - pos[t[0]]))
        for k in t:
            M[k] += a / 3.0
    return M


def shrink_rbc_volume(p, vol_final, n_it, damping=1e-2, max_damping=2.0):
    """shrink_rbc_volume (progs/common/calibrate_rbc_shape.h:13-72) on an
    oracle Particle of model 2: from the rest-shape volume, vol_desired
    ramps to vol_final over n_it relaxation steps, then 2 n_it + 1 steps
    in all; each step is relax_nodes (src/fsi/RBCParticle.hh:318-342):
    vel += F / M - damping vel, pos += vel, geometry update, new forces.
    A step whose area is NaN restarts the whole calibration from the
    starting state with twice the damping.  Returns the final damping; the
    particle ends in the calibrated state, its velocities in p.vel and
    forces in p.force.  The reference retries without end; here a damping
    past max_damping raises (the step is unstable for any damping above 2)."""
    vol_start = particle_update(p.tri, p.rest)["volume"]
    p0, prm0 = p.pos.copy(), p.params.copy()
    F0 = p.forces()
    vel0 = np.zeros_like(p.pos)
    while True:
        p.pos, p.params = p0.copy(), prm0.copy()
        g = particle_update(p.tri, p.pos)
        p.area, p.volume = g["area"], g["volume"]
        vel, F = vel0.copy(), F0.copy()
        ok = True
        for it in range(2 * n_it + 1):
            if it <= n_it:
                p.params[5] = vol_start + (float(it) / float(n_it)) * (vol_final - vol_start)
            M = node_masses(p.tri, p.pos)
            vel = vel + (F / M[:, None] - damping * vel)
            p.pos = p.pos + vel
            g = particle_update(p.tri, p.pos)
            p.area, p.volume, p.com = g["area"], g["volume"], g["com"]
            F = p.forces()
            if p.area != p.area:
                damping *= 2.0
                ok = False
                if damping > max_damping:
                    raise FloatingPointError("shrink_rbc_volume: diverges for every damping up to %g" % max_damping)
                break
        if ok:
            p.vel, p.force = vel, F
            return damping


def place_spheres(box, boundary, max_radius, batches, seed=None):
    """ParticlePositionInitializer (src/fsi/ParticlePositionInitializer.hh:
    18-86) with the C library's rand(): per batch (count, radius), each
    point is drawn as box corner + radius + rand()/RAND_MAX * (extent -
    2 radius) per axis (x, y, z), until it lies inside `boundary` (margin
    = radius) and overlaps no earlier sphere (distance^2 < (r1 + r2)^2);
    5e6 draws at most per point.  box = (x0, x1, y0, y1, z0, z1) in lattice
    cells.  Returns one (n, 3) array per batch, or None for a batch that
    could not be placed."""
    libc = ctypes.CDLL(None)
    if seed is not None:
        libc.srand(ctypes.c_uint(seed))
    RAND_MAX = 2147483647.0
    pts, out = [], []
    for count, radius in batches:
        if 2 * radius > 2 * max_radius:
            out.append(None)
            continue
        got = []
        for _ in range(count):
            placed = False
            for _ in range(int(5e6)):
                c = []
                for a in range(3):
                    lo, hi = box[2 * a], box[2 * a + 1]
                    c.append(lo + radius + (libc.rand() / RAND_MAX) * ((hi - lo) - 2 * radius))
                c = np.array(c)
                if boundary is not None and not lib.or_boundary_contains(ctypes.byref(boundary), _p(c), radius):
                    continue
                if any(((q - c) ** 2).sum() < (r + radius) ** 2 for q, r in pts):
                    continue
                placed = True
                break
            if not placed:
                break
            pts.append((c, radius))
            got.append(c)
        out.append(np.array(got) if len(got) == count else None)
    return out


# ---- rigid bodies (growth packing) -------------------------------------
lib.or_shape_inertia.restype = None
lib.or_rigid_update.restype = None
lib.or_rigid_move.restype = None
lib.or_rigid_move.argtypes = [ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p,
                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
lib.or_rigid_update.argtypes = [ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
lib.or_shape_inertia.argtypes = [ctypes.c_long, ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]


def shape_inertia(rest, tri):
    """Rest-shape moment of inertia about the mesh origin (3x3)."""
    rest = _c(rest, np.float64)
    tri = _c(tri, np.int32)
    out = np.empty(9)
    lib.or_shape_inertia(len(rest), len(tri), _p(tri), _p(rest), _p(out))
    return out.reshape(3, 3)


def rigid_state(com, quat, scale, density=1.0, damping=0.0, vel=(0, 0, 0), angvel=(0, 0, 0)):
    """The 16-double rigid state: com, velocity, angular velocity,
    quaternion (w, x, y, z), scale, density, damping."""
    return np.concatenate([np.asarray(com, float), np.asarray(vel, float), np.asarray(angvel, float),
                           np.asarray(quat, float), [scale, density, damping]])


def rigid_update(rest, state):
    rest = _c(rest, np.float64)
    out = np.empty_like(rest)
    lib.or_rigid_update(len(rest), _p(rest), _p(_c(state, np.float64)), _p(out))
    return out


def rigid_move(rest, inertia, volume, periodic, shape, state, pos, force):
    """One RigidParticle3D::move_vertices + periodic centre-of-mass remap;
    returns (state, pos) updated copies."""
    rest = _c(rest, np.float64)
    st = _c(state, np.float64).copy()
    p = _c(pos, np.float64).copy()
    lib.or_rigid_move(len(rest), _p(rest), _p(_c(inertia, np.float64).reshape(-1)), float(volume),
                      _p(_c(periodic, np.int32)), _p(_c(shape, np.int32)), _p(st), _p(p),
                      _p(_c(force, np.float64)))
    return st, p


def rigid_grow(rest, inertia, volume, shape, periodic, states, u, boundary, n_iter, initial_scale, every=10,
               record=()):
    """The growth loop of progs/common/create_particles.h:129-155 over rigid
    particles (states: (n, 16)): every `every` iterations the scale moves
    linearly from initial_scale to 1 and the vertices are regenerated; each
    iteration the collision handler's pair and wall forces, then every
    particle moves in order.  Returns (states, positions) after the last
    iteration and {k: (states, positions)} after each iteration in
    `record`."""
    states = np.array(states, dtype=np.float64)
    n = len(states)
    nv = len(rest)
    obj = np.repeat(np.arange(n, dtype=np.int32), nv)
    pos = np.concatenate([rigid_update(rest, s) for s in states])
    out = {}
    for it in range(n_iter + 1):
        if it % every == 0:
            scale = initial_scale + (1.0 - initial_scale) * (it / n_iter)
            states[:, 13] = scale
            pos = np.concatenate([rigid_update(rest, s) for s in states])
        force, _ = collision_forces(shape, periodic, u, pos, obj, wall=boundary)
        for k in range(n):
            sl = slice(k * nv, (k + 1) * nv)
            states[k], pos[sl] = rigid_move(rest, inertia, volume, periodic, shape, states[k], pos[sl], force[sl])
        if it in record:
            out[it] = (states.copy(), pos.copy())
    return states, pos, out
