"""The oracle's immersed-boundary restatement (oracle/ibm.c) against the
golden vectors written by the reference's own fsi classes
(oracle/ref/ref_fsi.cpp: ParticleShape, DeformableCapsuleParticle,
RBCParticle, DeformableParticle3D, RayTracer, Dirac), plus the closed forms
and invariants that pin the parts no reference build covers (interpolation
and spreading live in ImmersedBoundaryDynamics3D, which needs MPI).
CPU only."""
import numpy as np
import pytest

from conftest import load_golden
from biblbm_amd import meshes
from oracle import oracle_py as O

MESHES = ["sphere626", "ellipsoid414"]
KERNELS = [(0, "roma", 1.5), (1, "peskin", 2.0), (2, "tophat", 1.0)]


def g(mesh, name):
    return load_golden("fsi", mesh, name + ".npy")


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


# ------------------------------------------------------------ Dirac kernels
@pytest.mark.parametrize("k,name,h", KERNELS)
def test_dirac_matches_reference(k, name, h):
    x = load_golden("fsi", "dirac", "x.npy")
    ref = load_golden("fsi", "dirac", name + ".npy")
    inside = np.abs(x) <= h
    mine = np.where(inside, O.dirac(k, np.where(inside, x, 0.0)), 0.0)
    assert np.abs(mine - ref).max() <= 1e-15


def test_support_boxes_bit_exact():
    pts = load_golden("fsi", "dirac", "points.npy")
    base = load_golden("fsi", "dirac", "box_base.npy")
    for k, _, _ in KERNELS:
        assert np.array_equal(O.support_base(k, pts), base[:, k, :])


@pytest.mark.parametrize("k,name,h", KERNELS)
def test_dirac_moment_conditions(k, name, h):
    """sum_j phi(r - j) = 1 and sum_j (r - j) phi(r - j) = 0 over the stencil;
    sum_j phi^2 = 1/2 (Roma) and 3/8 (Peskin) for every offset."""
    rng = np.random.default_rng(3)
    S = {0: 3, 1: 4, 2: 2}[k]
    for X in rng.uniform(-5, 5, 200):
        b = int(O.support_base(k, np.array([X]))[0])
        r = np.arange(b, b + S) - X
        w = O.dirac(k, r)
        assert abs(w.sum() - 1.0) < 1e-14
        assert abs((r * w).sum()) < 1e-14
        if k == 0:
            assert abs((w * w).sum() - 0.5) < 1e-14
        if k == 1:
            assert abs((w * w).sum() - 0.375) < 1e-14


# --------------------------------------------------- rest shape / topology
@pytest.mark.parametrize("mesh", MESHES)
def test_topology_matches_reference(mesh):
    topo = O.build_topology(g(mesh, "tri"), g(mesh, "rest_pos"))
    assert np.array_equal(topo["link_i"], g(mesh, "link_i"))
    assert np.array_equal(topo["link_t"], g(mesh, "link_t"))
    assert np.abs(topo["link_len"] - g(mesh, "link_len")).max() < 1e-13
    assert np.array_equal(topo["pair_i"], g(mesh, "pair_i"))
    assert np.abs(topo["pair_cs0"] - g(mesh, "pair_cs0")).max() < 1e-13


@pytest.mark.parametrize("mesh", MESHES)
def test_euler_characteristic(mesh):
    tri, rest = g(mesh, "tri"), g(mesh, "rest_pos")
    topo = O.build_topology(tri, rest)
    V, E, F = len(rest), len(topo["link_i"]), len(tri)
    assert V - E + F == 2
    assert len(topo["pair_i"]) == E  # one dihedral pair per interior edge of a closed mesh


# ------------------------------------------------------- membrane forces
@pytest.mark.parametrize("mesh", MESHES)
@pytest.mark.parametrize("model", ["skalak", "neohookean"])
def test_capsule_forces_match_reference(mesh, model):
    F = O.capsule_forces(g(mesh, "tri"), g(mesh, "rest_pos"), g(mesh, "pos"), 0.02, 10.0, model)
    assert rel(F, g(mesh, "force_" + model)) < 1e-12


@pytest.mark.parametrize("mesh", MESHES)
def test_rbc_forces_match_reference(mesh):
    tri, pos = g(mesh, "tri"), g(mesh, "pos")
    topo = O.build_topology(tri, g(mesh, "rest_pos"))
    geo = O.particle_update(tri, pos)
    F = O.rbc_forces(tri, g(mesh, "tri_area"), topo, g(mesh, "shape_scalars")[0], g(mesh, "rbc_params"),
                     geo["area"], geo["volume"], pos)
    assert rel(F, g(mesh, "force_rbc")) < 1e-12


def test_planted_model_swap_is_caught():
    """A plausible mistake -- evaluating the NeoHookean law where Skalak is
    asked for -- fails the golden comparison."""
    mesh = MESHES[0]
    F = O.capsule_forces(g(mesh, "tri"), g(mesh, "rest_pos"), g(mesh, "pos"), 0.02, 10.0, "neohookean")
    assert rel(F, g(mesh, "force_skalak")) > 1e-3


@pytest.mark.parametrize("mesh", MESHES)
def test_forces_vanish_at_rest_and_are_free_of_net_force_and_torque(mesh):
    tri, rest, pos = g(mesh, "tri"), g(mesh, "rest_pos"), g(mesh, "pos")
    F0 = O.capsule_forces(tri, rest, rest, 0.02, 10.0, "skalak", bug_compatible=False)
    assert np.abs(F0).max() < 1e-14
    for F in (O.capsule_forces(tri, rest, pos, 0.02, 10.0, "skalak"),
              O.capsule_forces(tri, rest, pos, 0.02, 10.0, "neohookean")):
        scale = np.abs(F).max() * len(F)
        assert np.abs(F.sum(axis=0)).max() < 1e-12 * scale
        c = pos.mean(axis=0)
        assert np.abs(np.cross(pos - c, F).sum(axis=0)).max() < 1e-10 * scale


# ----------------------------------------------------- particle geometry
@pytest.mark.parametrize("mesh", MESHES)
def test_geometry_and_inertia_match_reference(mesh):
    tri, pos = g(mesh, "tri"), g(mesh, "pos")
    geo = O.particle_update(tri, pos)
    ref = g(mesh, "geometry")
    assert abs(geo["area"] - ref[0]) < 1e-10 * ref[0]
    assert abs(geo["volume"] - ref[1]) < 1e-10 * abs(ref[1])
    assert np.abs(geo["com"] - ref[2:5]).max() < 1e-12
    assert np.abs(geo["bbox"] - ref[5:]).max() == 0.0
    assert rel(O.moment_of_inertia(tri, pos), g(mesh, "inertia")) < 1e-12


def test_inertia_of_sphere_closed_form():
    v, t = meshes.icosphere(4, 10.0)
    I = O.moment_of_inertia(t, v)
    vol = O.particle_update(t, v)["volume"]
    # solid sphere about its centre: I = 2/5 m r^2 with m = volume (unit density)
    r_eff = (3 * vol / (4 * np.pi)) ** (1 / 3)
    assert abs(I[0, 0] / (0.4 * vol * r_eff ** 2) - 1) < 5e-3
    assert np.abs(I - np.diag(np.diag(I))).max() < 1e-8 * I[0, 0]


# ----------------------------------------------------------- voxelization
@pytest.mark.parametrize("mesh", MESHES)
def test_ray_hits_match_reference(mesh):
    h = O.ray_hits_z(g(mesh, "tri"), g(mesh, "pos"))
    ref = g(mesh, "ray_hits_z")
    assert h.shape == ref.shape
    assert np.abs(h - ref).max() < 1e-12


def test_voxel_fill_counts_sphere_volume():
    """Centre off the lattice planes so that no z-ray passes through a mesh
    vertex or edge; every lattice point then matches the inside test of the
    sphere except within the chord gap of the polyhedron."""
    c = np.array([16.37, 15.71, 16.13])
    r = 9.3
    v, t = meshes.icosphere(4, r)
    v = v + c
    phase = O.voxel_fill((32, 32, 32), (0, 0, 0), O.ray_hits_z(t, v))
    x, y, z = np.meshgrid(*[np.arange(32.0)] * 3, indexing="ij")
    d = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)
    wrong = (phase > 0) != (d < r)
    assert np.all(np.abs(d[wrong] - r) < 0.02)
    assert abs(phase.sum() / O.particle_update(t, v)["volume"] - 1) < 5e-3


# ------------------------------------------- interpolation and spreading
@pytest.mark.parametrize("k,name,h", KERNELS)
def test_interpolation_exact_for_linear_fields(k, name, h):
    rng = np.random.default_rng(5)
    shape = (16, 14, 12)
    x, y, z = np.meshgrid(*[np.arange(n, dtype=float) for n in shape], indexing="ij")
    A = rng.normal(size=(3, 3))
    b = rng.normal(size=3)
    u = np.stack([A[d, 0] * x + A[d, 1] * y + A[d, 2] * z + b[d] for d in range(3)], axis=-1)
    pos = rng.uniform(3, 9, size=(50, 3))
    U = O.interpolate(k, u, (0, 0, 0), pos)
    assert np.abs(U - (pos @ A.T + b)).max() < 1e-12


@pytest.mark.parametrize("k,name,h", KERNELS)
def test_spread_conserves_force_and_is_adjoint(k, name, h):
    rng = np.random.default_rng(6)
    shape = (12, 10, 11)
    pos = rng.uniform(0, 12, size=(40, 3))
    Fv = rng.normal(size=(40, 3))
    Fg = O.spread(k, shape, (1, 1, 1), pos, Fv)
    assert np.abs(Fg.reshape(-1, 3).sum(axis=0) - Fv.sum(axis=0)).max() < 1e-12
    u = rng.normal(size=shape + (3,))
    lhs = (O.interpolate(k, u, (1, 1, 1), pos) * Fv).sum()
    rhs = (u * Fg).sum()
    assert abs(lhs - rhs) < 1e-11 * abs(rhs)


def test_periodic_wrap_equals_shifted_interpolation():
    rng = np.random.default_rng(7)
    shape = (10, 10, 10)
    u = rng.normal(size=shape + (3,))
    pos = rng.uniform(0, 10, size=(30, 3))
    a = O.interpolate(1, u, (1, 1, 1), pos)
    b = O.interpolate(1, np.roll(u, 3, axis=0), (1, 1, 1), pos + [3, 0, 0])
    assert np.abs(a - b).max() < 1e-13


# ------------------------------------------------------------ mesh builders
@pytest.mark.parametrize("level,nv,nt", [(2, 162, 320), (3, 642, 1280), (4, 2562, 5120)])
def test_icosphere_counts_and_orientation(level, nv, nt):
    v, t = meshes.icosphere(level, 3.0)
    assert v.shape == (nv, 3) and t.shape == (nt, 3)
    assert np.allclose(np.linalg.norm(v, axis=1), 3.0)
    assert meshes.signed_volume(v, t) > 0


def test_gmsh_round_trip(tmp_path):
    v, t = meshes.biconcave(2, 4.0)
    p = tmp_path / "rbc.msh"
    meshes.write_gmsh(str(p), v, t)
    v2, t2 = meshes.read_gmsh(str(p))
    assert np.allclose(v, v2) and np.array_equal(t, t2)
