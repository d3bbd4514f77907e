"""Composed coupled steps on the GPU (through the C-ABI) against closed forms
that only the reference's step order satisfies, and against the oracle on
reduced forms of the capsule-in-shear, Taylor-Green and suspension
workloads; size-independent properties of the full 400-cell suspension.

Bars (FP64 storage): vertex positions, density and velocity within
relative L2 1e-10 of the oracle (positions also 1e-8 relative to the
displacement), Taylor deformation index within 1e-8, closed forms within
1e-12 (drift) and 1e-10 (momentum budget) relative; reduced storage:
velocity relative L2 against the FP64 run stated per case.
"""
import numpy as np
import pytest

import test_oracle_coupled as T
from oracle import oracle_py as O

pytestmark = pytest.mark.gpu

PER = (1, 1, 1)


def oracle_particles(ib):
    """Oracle particles with the device's current vertex positions and the
    shapes, models and parameters the boundary was given."""
    pos = ib.positions()
    out, k = [], 0
    for sid, model, params, bug in ib.specs:
        rest, tri = ib.meshes[sid]
        out.append(O.Particle(rest, tri, model, params, pos[k:k + len(rest)], bug))
        k += len(rest)
    return out


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def assert_matches(L, ib, f, force, parts, start, mask=None, wall_u=None, tag=""):
    from biblbm_amd.ibm import taylor_deformation
    ref = np.concatenate([p.pos for p in parts])
    pos = ib.positions()
    e_pos, e_disp = rel(pos, ref), np.linalg.norm(pos - ref) / np.linalg.norm(ref - start)
    rho, u = L.moments()
    rho_ref, u_ref = O.lbm_moments(f, force, mask, 1, wall_u)
    fluid = np.ones(rho.shape, bool) if mask is None else (mask == 0) | (mask == 3)
    e_rho, e_u = rel(rho[fluid], rho_ref[fluid]), rel(u[fluid], u_ref[fluid])
    dD = max(abs(taylor_deformation(ib.observables(k)["inertia"]) -
                 taylor_deformation(O.moment_of_inertia(p.tri, p.pos))) for k, p in enumerate(parts))
    print(tag, "positions", e_pos, "displacement", e_disp, "rho", e_rho, "u", e_u, "D", dD)
    assert e_pos <= 1e-10 and e_disp <= 1e-8
    assert e_rho <= 1e-10 and e_u <= 1e-10
    assert dD <= 1e-8


# ----------------------------------------------------- closed forms
def _drift_box(f0, F0=None, body=None):
    from biblbm_amd.lattice import Lattice
    from biblbm_amd.ibm import ImmersedBoundary
    v, t, start = T.drift_capsule()
    L = Lattice(*T.DRIFT_SHAPE, periodic=PER, storage="f64", omega=T.OMEGA)
    L.put_populations(f0)
    if F0 is not None:
        L.set_force_field(F0)
    if body is not None:
        L.set_body_force(body)
    ib = ImmersedBoundary(L, "peskin")
    sid = ib.add_shape(v, t)
    ib.add_particle(sid, 0, [0.02, 10.0], start)
    ib.init()
    return L, ib, start


def test_drift_with_per_step_body_force(device):
    """Per-step body forces g_n: the rest capsule moves by
    sum_n (u0 + sum_{k<n} g_k + g_{n-1}/2), which holds only if the staged
    velocity carries the last collision's force and the collision the new
    one."""
    f0, F0 = T.drift_start()
    L, ib, start = _drift_box(f0, F0)
    bodies = T.drift_bodies()
    for g in bodies:
        L.set_body_force(g)
        ib.coupled_steps(1, use_graph=False)
    disp = T.expected_drift(T.DRIFT_U0, T.DRIFT_F_INIT, bodies)
    err = np.abs(ib.positions() - (start + disp)).max()
    print("drift error", err, "displacement", np.abs(disp).max())
    assert err <= 1e-12 * np.abs(disp).max()


def test_drift_constant_body_force_graph(device):
    """Constant body force g from the first step, graph-replayed runs
    (a 32-step run and pairs): displacement N u0 + g N^2 / 2."""
    g = np.array([1.5e-4, -1.0e-4, 0.5e-4])
    f0, _ = T.drift_start()
    L, ib, start = _drift_box(f0, None, g)
    n = 40
    ib.coupled_steps(n, use_graph=True)
    disp = n * T.DRIFT_U0 + g * n * n / 2.0
    assert np.allclose(disp, T.expected_drift(T.DRIFT_U0, g, [g] * n), rtol=1e-14, atol=0)
    err = np.abs(ib.positions() - (start + disp)).max()
    print("drift error", err, "displacement", np.abs(disp).max())
    assert err <= 1e-12 * np.abs(disp).max()


def test_momentum_budget(device):
    """Per step: the total momentum changes by sum_x F(x)(1 + omega(rho-1)/2)
    for the force field the lattice reports after the step, whose sum is
    N g + sum_k F_k of the step's vertex forces; those equal the oracle's at
    the positions before the step, and the vertex velocities equal the
    interpolation of j/rho + F_prev/2 at those positions."""
    from biblbm_amd.lattice import Lattice
    from biblbm_amd.ibm import ImmersedBoundary
    from biblbm_amd.potentials import MorsePotential
    from biblbm_amd.walls import Boundary
    f, force, v, t, start, wall = T.budget_setup()
    L = Lattice(*T.BUDGET_SHAPE, periodic=PER, storage="f64", omega=T.OMEGA)
    L.put_populations(f)
    L.set_force_field(force)
    L.set_body_force(T.BUDGET_BODY)
    ib = ImmersedBoundary(L, "peskin")
    sid = ib.add_shape(v, t)
    ib.add_particle(sid, 0, [0.02, 10.0], start)
    ib.init()
    ib.add_wall_interaction(Boundary.plates(1.0, 18.0), MorsePotential(4.0, 1.5, 2e-3))
    p_or = O.Particle(v, t, 0, [0.02, 10.0], start)
    for step in range(T.BUDGET_STEPS + 2):
        f_before, F_prev, x_n = L.get_populations(), L.force(), ib.positions()
        ib.coupled_steps(1, use_graph=False)
        f_after, F_used, Fk = L.get_populations(), L.force(), ib.forces()
        T.budget_check(f_before, f_after, F_used, Fk, T.OMEGA, T.BUDGET_SHAPE, T.BUDGET_BODY)
        p_or.pos = x_n.copy()
        fk_ref = O.vertex_forces([p_or], T.BUDGET_SHAPE, PER, [wall])[0]
        assert np.abs(Fk - fk_ref).max() <= 1e-12 * np.abs(fk_ref).max()
        _, u_st = O.lbm_moments(f_before, F_prev)
        vel_ref = O.interpolate("peskin", u_st, PER, x_n)
        assert np.abs(ib.velocities() - vel_ref).max() <= 1e-12 * np.abs(vel_ref).max()


def test_voxel_fill_across_periodic_z_face(device):
    """The device's phase field of a sphere crossing the periodic z face
    equals the brute-force statement of the reference's fill rule
    (remapped hits, sorted, parity fill; inverted on the columns whose hits
    straddle the face)."""
    from biblbm_amd.lattice import Lattice
    from biblbm_amd.ibm import ImmersedBoundary
    pos, t = T.straddling_sphere()
    L = Lattice(*T.STRADDLE_SHAPE, periodic=PER, storage="f64", omega=1 / 0.8, omega_in=1 / 2.0)
    ib = ImmersedBoundary(L, "peskin")
    sid = ib.add_shape(pos - T.STRADDLE_CENTRE, t)
    ib.add_particle(sid, 1, [1e-3, 1.0], pos)
    ib.init()
    before = ib.phase()
    ib.step()
    p = ib.positions()
    ph = ib.phase()
    assert np.array_equal(ph, T.brute_force_phase(T.STRADDLE_SHAPE, t, p, old=before))
    hits = O.ray_hits_z(t, p)
    ref = O.voxel_fill(T.STRADDLE_SHAPE, PER, hits)
    assert np.array_equal(ph, ref)


# ------------------------------------------- reduced workloads vs oracle
def test_capsule_in_shear_matches_oracle(device):
    """capsule_in_shear at 24^3 (velocity walls at y = 0, 23 moving at
    -/+0.01, tau 1, Skalak capsule of 642 vertices, Ca 0.1, Peskin),
    graph-replayed, against the oracle's composed steps."""
    from biblbm_amd import workloads
    W = workloads.capsule_in_shear("f64", n=24, radius=5.0, level=3)
    L, ib = W.lattice, W.ibm
    n = 24
    f = L.get_populations()
    parts = oracle_particles(ib)
    start = np.concatenate([p.pos for p in parts])
    mask = np.zeros((n, n, n), np.uint8)
    mask[:, 0, :] = mask[:, -1, :] = 3
    wall_u = np.array([-0.01, 0, 0, 0.01, 0, 0])
    steps = 8
    ib.coupled_steps(steps, use_graph=True)
    force = np.zeros((n, n, n, 3))
    om = 1 / W.params["tau"]
    for _ in range(steps):
        f, force, _ = O.coupled_step(f, force, np.zeros((n, n, n)), mask, (1, 0, 1), om, om, (0, 0, 0), parts,
                                     "peskin", 1, wall_u)
    assert_matches(L, ib, f, force, parts, start, mask, wall_u, "capsule_in_shear")
    ib.close()
    L.close()


def test_taylor_green_reduced_matches_oracle(device):
    """The precision workload reduced to 48^3 with 2 Skalak capsules (642
    vertices, radius 5): FP64 device steps against the oracle."""
    from biblbm_amd import workloads
    W = workloads.taylor_green("f64", n=48, n_caps=2, radius=5.0, level=3)
    L, ib = W.lattice, W.ibm
    f = L.get_populations()
    parts = oracle_particles(ib)
    start = np.concatenate([p.pos for p in parts])
    shape = L.shape
    steps = 6
    ib.coupled_steps(steps, use_graph=True)
    force = np.zeros(shape + (3,))
    om = 1 / W.params["tau"]
    for _ in range(steps):
        f, force, _ = O.coupled_step(f, force, np.zeros(shape), np.zeros(shape, np.uint8), PER, om, om, (0, 0, 0),
                                     parts, "peskin")
    assert_matches(L, ib, f, force, parts, start, tag="taylor_green 48^3")
    ib.close()
    L.close()


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f16s", 1e-3)])
def test_taylor_green_reduced_storage_error(device, storage, tol):
    """FP32 and FP16-shifted storage (FP32 collision) of the reduced
    Taylor-Green workload against its FP64 run after 60 steps: velocity
    relative L2 within the stated bound, vertex displacement within 10x."""
    from biblbm_amd import workloads
    runs = {}
    for s in ("f64", storage):
        W = workloads.taylor_green(s, n=48, n_caps=2, radius=5.0, level=3)
        p0 = W.ibm.positions()
        W.ibm.coupled_steps(60, use_graph=True)
        _, u = W.lattice.moments()
        runs[s] = (u, W.ibm.positions(), p0)
        W.ibm.close()
        W.lattice.close()
    u64, x64, p0 = runs["f64"]
    u, x, _ = runs[storage]
    e_u = rel(u, u64)
    e_x = np.linalg.norm(x - x64) / np.linalg.norm(x64 - p0)
    print(storage, "velocity rel L2", e_u, "displacement rel", e_x)
    assert e_u <= tol
    assert e_x <= 10 * tol


def _small_suspension(storage="f64"):
    from biblbm_amd import workloads
    return workloads.suspension(storage, nx=64, ny=40, nz=48, n642=2, n2562=1, n162=2, cell_radius=8.0, seed=3)


@pytest.mark.parametrize("fused", [-1, 1])
def test_suspension_reduced_matches_oracle(device, fused):
    """The 400-cell workload reduced to 64 x 40 x 48 with the same mix of
    meshes (2 x 642 and 1 x 2562 red cells with Skalak + bending + area +
    volume, 2 platelet capsules of 162 vertices), bounce-back walls and
    body force; staging carried by the collision forced on (1) or left to
    the vertex count (-1: separate pass here), against the oracle."""
    W = _small_suspension()
    L, ib = W.lattice, W.ibm
    ib.set_fused_staging(fused)
    shape = L.shape
    f = L.get_populations()
    parts = oracle_particles(ib)
    assert sorted({len(p.pos) for p in parts}) == [162, 642, 2562]
    start = np.concatenate([p.pos for p in parts])
    mask = np.zeros(shape, np.uint8)
    mask[:, 0, :] = mask[:, -1, :] = 1
    g = (W.params["g"], 0.0, 0.0)
    steps = 5
    ib.coupled_steps(steps, use_graph=True)
    force = np.broadcast_to(np.array(g), shape + (3,)).copy()
    om = 1 / W.params["tau"]
    for _ in range(steps):
        f, force, _ = O.coupled_step(f, force, np.zeros(shape), mask, (1, 0, 1), om, om, g, parts, "peskin")
    assert_matches(L, ib, f, force, parts, start, mask, tag="suspension reduced fused=%d" % fused)
    ib.close()
    L.close()


def test_suspension_full_size_properties(device):
    """The full 512 x 128 x 128 suspension (400 cells, 400 800 vertices) in
    FP64: the population sum over every cell (mass) is kept to 1e-12
    relative over 12 graph-replayed steps, and two runs from the same
    seeded start are bitwise equal."""
    from biblbm_amd import workloads
    runs = []
    for k in range(2):
        W = workloads.suspension("f64")
        if k == 0:
            m0 = W.lattice.get_populations().sum() + W.cells
        W.ibm.coupled_steps(12, use_graph=True)
        f = W.lattice.get_populations()
        runs.append((f, W.ibm.positions()))
        if k == 0:
            m1 = f.sum() + W.cells
            print("mass", m0, m1, "relative change", abs(m1 - m0) / m0)
            assert abs(m1 - m0) <= 1e-12 * m0
        W.ibm.close()
        W.lattice.close()
        del W
    assert np.array_equal(runs[0][1], runs[1][1])
    assert np.array_equal(runs[0][0], runs[1][0])
