"""Builders for the benchmark / parity configurations (seeded, synthetic):

  capsule_in_shear   64^3, periodic x/z, regularized velocity walls at y = 0
                     and 63 moving at -/+0.01, tau 1.0, one Skalak capsule
                     (642-vertex icosphere, radius 8, Ca 0.1, C 10), Peskin.
  rbc_in_pipe        pipe radius 32, length 257 (bbox 257 x 67 x 67,
                     staircase bounce-back, or with second_order the Guo
                     curved wall and no dynamics beyond it, the reference
                     prog's default), body force for centreline
                     velocity 0.02, tau 0.9; one Evans-Fung red cell on the
                     2562-vertex icosphere, radius 12: Skalak in-plane +
                     bending kb = 0.02 G R^2 + global area and volume
                     constraints; Peskin.
  suspension         512 x 128 x 128 channel (bounce-back planes y = 0, 127),
                     body force, tau 1.0; 300 red cells of 642 vertices, 80 of
                     2562 (radius 13.5), 20 platelet ellipsoids of 162,
                     rejection-sampled with random orientations (17.7 %
                     volume fraction: random sequential addition of these
                     discs jams a little below 20 %).
  taylor_green       384^3 periodic Taylor-Green (u0 0.03, tau 0.8) with 16
                     Skalak capsules (642 vertices) at seeded positions.
  shear_ensemble     64 independent shear boxes, sides drawn from
                     {32, 40, 48, 56, 64}, velocity walls y = 0 / s-1, tau
                     uniform in [0.7, 1.2], one 642-vertex biconcave cell
                     (radius 0.16 s, shear modulus 0.002) per box, the wall
                     speed setting Ca log-uniform in [0.05, 0.5]; Peskin.

The red-cell constraint parameters follow the reference's create_rbc_params
(progs/common/calibrate_rbc_shape.h:75-112): area modulus split 0.8 / 0.2
between global and local, k_volume = 1000 G / R.
"""
from dataclasses import dataclass, field

import numpy as np

from . import meshes
from .ibm import ImmersedBoundary, CAPSULE_SKALAK, SKALAK_RBC
from .lattice import Lattice, BOUNCE_BACK, VELOCITY_WALL
from .potentials import MorsePotential
from .walls import Boundary, GuoWall, outside_mask


@dataclass
class Workload:
    name: str
    lattice: Lattice
    ibm: ImmersedBoundary
    cells: int
    fluid_cells: int
    vertices: int
    params: dict = field(default_factory=dict)
    guo: object = None  # a Guo curved wall the lattice runs (kept alive with it)


def rbc_constraint_params(G, radius, volume, K_area_over_G=20.0, kb=0.0):
    delta = K_area_over_G * G - 2.0 * G
    return dict(k_area_global=0.8 * delta, k_area_local=0.2 * delta,
                k_volume=1000.0 * G / radius, vol_desired=volume, k_bend=kb)


def skalak_rbc_params(G, C, radius, volume, kb):
    c = rbc_constraint_params(G, radius, volume, kb=kb)
    return [G, C, c["k_area_global"], c["k_area_local"], c["k_volume"], c["vol_desired"], c["k_bend"]]


def _shape_volume(v, t):
    return meshes.signed_volume(v, t)


def capsule_in_shear(storage="f64", n=64, tau=1.0, wall_u=0.01, radius=8.0, level=3, Ca=0.1, C=10.0,
                     kernel="peskin", model=CAPSULE_SKALAK, bug_compatible=True):
    L = Lattice(n, n, n, periodic=(1, 0, 1), storage=storage, omega=1 / tau)
    L.set_velocity_walls(1, [-wall_u, 0, 0], [wall_u, 0, 0])
    mask = np.zeros((n, n, n), np.uint8)
    mask[:, 0, :] = mask[:, -1, :] = VELOCITY_WALL
    L.set_mask(mask)
    y = np.arange(n)
    ux = wall_u * (2.0 * y / (n - 1) - 1.0)
    u = np.zeros((n, n, n, 3))
    u[..., 0] = ux[None, :, None]
    L.init_equilibrium(None, u)
    L.initialize()
    nu = (tau - 0.5) / 3.0
    gamma = 2 * wall_u / (n - 1)
    G = nu * gamma * radius / Ca
    v, t = meshes.icosphere(level, radius)
    ib = ImmersedBoundary(L, kernel)
    sid = ib.add_shape(v, t)
    centre = np.array([n / 2.0, (n - 1) / 2.0, n / 2.0])
    ib.add_particle(sid, model, [G, C], v + centre, bug_compatible)
    ib.init()
    return Workload("capsule_in_shear", L, ib, n ** 3, n ** 3, len(v),
                    dict(n=n, tau=tau, wall_u=wall_u, G=G, C=C, Ca=Ca, radius=radius, kernel=kernel,
                         shear_rate=gamma))


def pipe_mask(R, length):
    ny = nz = 2 * R + 3
    c = (ny - 1) / 2.0
    yy, zz = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
    r2 = (yy - c) ** 2 + (zz - c) ** 2
    wall2d = r2 > R * R
    mask = np.broadcast_to(np.where(wall2d, BOUNCE_BACK, 0).astype(np.uint8), (length, ny, nz)).copy()
    return mask, r2, c


def rbc_in_pipe(storage="f64", R=32, length=257, u_c=0.02, tau=0.9, level=4, cell_radius=12.0, Ca=0.1,
                C=10.0, kernel="peskin", second_order=False):
    """second_order: the reference prog's default wall
    (progs/rbc_in_pipe/rbc_in_pipe.cpp:302-318): Guo curved wall on the pipe
    surface, no dynamics beyond it (margin 1); otherwise bounce-back on every
    cell outside the pipe."""
    mask, r2, c = pipe_mask(R, length)
    ny = nz = mask.shape[1]
    L = Lattice(length, ny, nz, periodic=(1, 0, 0), storage=storage, omega=1 / tau)
    guo = None
    if second_order:
        b = Boundary.pipe(c, c, R, float(length))
        guo = GuoWall(L, b)
        guo.insert()
        mask = outside_mask(b, mask.shape, 1.0)
    L.set_mask(mask)
    nu = (tau - 0.5) / 3.0
    g = 4.0 * nu * u_c / (R * R)
    L.set_body_force([g, 0, 0])
    prof = np.where(r2 <= R * R, u_c * (1.0 - r2 / (R * R)), 0.0)
    u = np.zeros((length, ny, nz, 3))
    u[..., 0] = prof[None, :, :]
    L.init_equilibrium(None, u)
    if guo is not None:
        L.initialize()
    G = nu * (u_c / R) * cell_radius / Ca
    kb = 0.02 * G * cell_radius ** 2
    k_bend = 2.0 * kb / np.sqrt(3.0)
    v, t = meshes.biconcave(level, cell_radius)
    vol = _shape_volume(v, t)
    ib = ImmersedBoundary(L, kernel)
    sid = ib.add_shape(v, t)
    rot = meshes.rotation_matrix([0, 1, 0], np.pi / 2)   # disc normal along the pipe axis
    pos = v @ rot.T + [length / 2.0, c, c]
    ib.add_particle(sid, SKALAK_RBC, skalak_rbc_params(G, C, cell_radius, vol, k_bend), pos)
    ib.init()
    fluid = int((mask == 0).sum())
    return Workload("rbc_in_pipe", L, ib, mask.size, fluid, len(v),
                    dict(R=R, length=length, u_c=u_c, tau=tau, G=G, kb=kb, k_bend=k_bend, g=g,
                         cell_radius=cell_radius, kernel=kernel, second_order=second_order,
                         wall_nodes=0 if guo is None else guo.n), guo)


def _solid_samples(v):
    c = v.mean(0)
    return np.concatenate([c + s * (v - c) for s in (1.0, 0.75, 0.5, 0.25)] + [c[None]])


def place_cells(shape, kinds, seed, periodic=(1, 0, 1), margin=1.5, y_margin=2.0, max_tries=4000):
    """Sequential rejection sampling of rigidly rotated meshes: a candidate is
    accepted when none of its solid sample points falls within `margin` of an
    occupied voxel and it keeps `y_margin` from the non-periodic faces."""
    rng = np.random.default_rng(seed)
    occ = np.zeros(shape, bool)
    out = []
    rad = int(np.ceil(margin))
    offs = np.array([(i, j, k) for i in range(-rad, rad + 1) for j in range(-rad, rad + 1)
                     for k in range(-rad, rad + 1) if i * i + j * j + k * k <= margin * margin + 1e-9])
    for kind, (v, t) in kinds:
        for attempt in range(max_tries):
            axis = rng.normal(size=3)
            R = meshes.rotation_matrix(axis, rng.uniform(0, 2 * np.pi))
            vr = v @ R.T
            lo, hi = vr.min(0), vr.max(0)
            centre = np.empty(3)
            ok = True
            for d in range(3):
                if periodic[d]:
                    centre[d] = rng.uniform(0, shape[d])
                else:
                    a, b = -lo[d] + y_margin, shape[d] - 1 - hi[d] - y_margin
                    if b <= a:
                        ok = False
                        break
                    centre[d] = rng.uniform(a, b)
            if not ok:
                continue
            pts = _solid_samples(vr + centre)
            idx = np.rint(pts).astype(int)
            for d in range(3):
                idx[:, d] = idx[:, d] % shape[d] if periodic[d] else np.clip(idx[:, d], 0, shape[d] - 1)
            if occ[idx[:, 0], idx[:, 1], idx[:, 2]].any():
                continue
            dil = (idx[:, None, :] + offs[None, :, :]).reshape(-1, 3)
            for d in range(3):
                dil[:, d] = dil[:, d] % shape[d] if periodic[d] else np.clip(dil[:, d], 0, shape[d] - 1)
            occ[dil[:, 0], dil[:, 1], dil[:, 2]] = True
            out.append((kind, vr + centre))
            break
        else:
            raise RuntimeError("could not place cell %d of kind %s" % (len(out), kind))
    return out


def suspension(storage="f32", nx=512, ny=128, nz=128, tau=1.0, u_c=0.02, n642=300, n2562=80, n162=20,
               cell_radius=13.5, Ca=0.1, C=10.0, kernel="peskin", seed=7, interactions=False):
    L = Lattice(nx, ny, nz, periodic=(1, 0, 1), storage=storage, omega=1 / tau)
    mask = np.zeros((nx, ny, nz), np.uint8)
    mask[:, 0, :] = mask[:, -1, :] = BOUNCE_BACK
    L.set_mask(mask)
    nu = (tau - 0.5) / 3.0
    H = (ny - 2) / 2.0
    g = 2.0 * nu * u_c / (H * H)
    L.set_body_force([g, 0, 0])
    yc = (ny - 1) / 2.0
    prof = np.clip(u_c * (1.0 - ((np.arange(ny) - yc) / H) ** 2), 0, None)
    prof[0] = prof[-1] = 0
    u = np.zeros((nx, ny, nz, 3))
    u[..., 0] = prof[None, :, None]
    L.init_equilibrium(None, u)
    G = nu * (u_c / H) * cell_radius / Ca
    k_bend = 2.0 * 0.02 * G * cell_radius ** 2 / np.sqrt(3.0)
    m642 = meshes.biconcave(3, cell_radius)
    m2562 = meshes.biconcave(4, cell_radius)
    m162 = meshes.ellipsoid(2, (3.0, 3.0, 1.2))
    kinds = [("rbc2562", m2562)] * n2562 + [("rbc642", m642)] * n642 + [("platelet", m162)] * n162
    placed = place_cells((nx, ny, nz), kinds, seed, margin=1.0, max_tries=20000)
    ib = ImmersedBoundary(L, kernel)
    sids = {"rbc642": ib.add_shape(*m642), "rbc2562": ib.add_shape(*m2562), "platelet": ib.add_shape(*m162)}
    vols = {k: _shape_volume(*m) for k, m in (("rbc642", m642), ("rbc2562", m2562), ("platelet", m162))}
    nv = 0
    for kind, pos in placed:
        if kind == "platelet":
            ib.add_particle(sids[kind], CAPSULE_SKALAK, [20 * G, C], pos)
        else:
            ib.add_particle(sids[kind], SKALAK_RBC, skalak_rbc_params(G, C, cell_radius, vols[kind], k_bend), pos)
        nv += len(pos)
    ib.init()
    if interactions:
        # the rbc_in_pipe prog's input.xml values: Morse wall interaction on
        # the bounce-back half-way planes, Morse particle-particle force
        ib.add_wall_interaction(Boundary.plates(0.5, ny - 1.5), MorsePotential(2.0, 2.0, 1e-4))
        ib.add_pp_force(MorsePotential(1.0, 1.0, 1e-3))
    vf = (n642 * vols["rbc642"] + n2562 * vols["rbc2562"] + n162 * vols["platelet"]) / (nx * (ny - 2) * nz)
    return Workload("suspension", L, ib, nx * ny * nz, nx * (ny - 2) * nz, nv,
                    dict(nx=nx, ny=ny, nz=nz, tau=tau, G=G, g=g, cells=len(placed), volume_fraction=vf,
                         kernel=kernel, seed=seed))


def suspension_interactions(storage="f32", **kw):
    """The suspension with the wall-interaction decorator and the
    particle-particle force switched on."""
    return suspension(storage, interactions=True, **kw)


def taylor_green_fields(n, u0):
    """Density and velocity of the Taylor-Green start on an n^3 periodic box
    (k = 2 pi / n; rho from the Taylor-Green pressure, c_s^2 = 1/3)."""
    k = 2 * np.pi / n
    x = np.arange(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij", sparse=True)
    u = np.zeros((n, n, n, 3))
    u[..., 0] = u0 * np.sin(k * X) * np.cos(k * Y) * np.cos(k * Z)
    u[..., 1] = -u0 * np.cos(k * X) * np.sin(k * Y) * np.cos(k * Z)
    rho = 1.0 + 3.0 * u0 ** 2 / 16.0 * (np.cos(2 * k * X) + np.cos(2 * k * Y)) * (np.cos(2 * k * Z) + 2)
    return np.broadcast_to(rho, (n, n, n)), u


def taylor_green(storage="f64", n=384, u0=0.03, tau=0.8, n_caps=16, radius=8.0, level=3, Ca=0.1, C=10.0,
                 kernel="peskin", seed=11):
    L = Lattice(n, n, n, storage=storage, omega=1 / tau)
    k = 2 * np.pi / n
    rho, u = taylor_green_fields(n, u0)
    L.init_equilibrium(rho, u)
    del u, rho
    nu = (tau - 0.5) / 3.0
    G = nu * (u0 * k) * radius / Ca
    v, t = meshes.icosphere(level, radius)
    placed = place_cells((n, n, n), [("capsule", (v, t))] * n_caps, seed, periodic=(1, 1, 1), margin=3.0)
    ib = ImmersedBoundary(L, kernel)
    sid = ib.add_shape(v, t)
    for _, pos in placed:
        ib.add_particle(sid, CAPSULE_SKALAK, [G, C], pos)
    ib.init()
    return Workload("taylor_green", L, ib, n ** 3, n ** 3, n_caps * len(v),
                    dict(n=n, u0=u0, tau=tau, G=G, kernel=kernel, seed=seed))


def shear_box(storage, s, tau, Ca, G=0.002, level=3, C=10.0, kernel="peskin", radius=None, seed=0):
    """One rbc_in_shear-style box: s^3 lattice, periodic x / z, regularized
    velocity walls at y = 0 and s - 1, linear shear start, one biconcave
    cell (Skalak + bending + area + volume, shear modulus G) at the centre
    with a seeded orientation.  The shear rate sets the capillary number:
    gamma = Ca G / (mu R), walls at -/+ gamma (s - 1) / 2 (below 0.05 for
    the ensemble's draws; a fixed G keeps the explicit membrane update
    stable for every draw)."""
    R = 0.16 * s if radius is None else radius
    mu = (tau - 0.5) / 3.0
    gamma = Ca * G / (mu * R)
    wall_u = gamma * (s - 1) / 2.0
    L = Lattice(s, s, s, periodic=(1, 0, 1), storage=storage, omega=1 / tau)
    L.set_velocity_walls(1, [-wall_u, 0, 0], [wall_u, 0, 0])
    mask = np.zeros((s, s, s), np.uint8)
    mask[:, 0, :] = mask[:, -1, :] = VELOCITY_WALL
    L.set_mask(mask)
    y = np.arange(s)
    u = np.zeros((s, s, s, 3))
    u[..., 0] = (wall_u * (2.0 * y / (s - 1) - 1.0))[None, :, None]
    L.init_equilibrium(None, u)
    L.initialize()
    v, t = meshes.biconcave(level, R)
    rng = np.random.default_rng(seed)
    axis = rng.normal(size=3)
    rot = meshes.rotation_matrix(axis / np.linalg.norm(axis), rng.uniform(0, np.pi))
    ib = ImmersedBoundary(L, kernel)
    sid = ib.add_shape(v, t)
    ib.add_particle(sid, SKALAK_RBC, skalak_rbc_params(G, C, R, _shape_volume(v, t), 0.02 * G * R * R),
                    v @ rot.T + [s / 2.0, (s - 1) / 2.0, s / 2.0])
    ib.init()
    return Workload("shear_box", L, ib, s ** 3, s ** 3, len(v),
                    dict(s=s, tau=tau, Ca=Ca, G=G, radius=R, wall_u=wall_u, shear_rate=gamma))


def shear_ensemble_params(n_boxes=64, seed=11, sides=(32, 40, 48, 56, 64)):
    """The seeded draws of the batched-ensemble configuration: per box
    (side, tau, capillary number, box seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(n_boxes):
        s = int(rng.choice(sides))
        tau = float(rng.uniform(0.7, 1.2))
        Ca = float(np.exp(rng.uniform(np.log(0.05), np.log(0.5))))
        out.append((s, tau, Ca, seed * 1000 + b))
    return out


def shear_ensemble(storage="f64", n_boxes=64, seed=11, sides=(32, 40, 48, 56, 64), **kw):
    """The boxes of the batched-ensemble configuration (seeded draws)."""
    return [shear_box(storage, s, tau, Ca, seed=bs, **kw)
            for s, tau, Ca, bs in shear_ensemble_params(n_boxes, seed, sides)]
