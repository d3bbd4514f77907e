"""Curved walls on the GPU (libbiblbm.so through the C-ABI) against the
reference's golden vectors and the FP64 oracle.

Bar: the wall-node list equals the reference's exactly (it is built on the
host with the reference's expressions); FP64 populations within 1e-12 of
the reference's (relative to their largest magnitude) on fluid and wall
cells after one step and after N steps; FP32 / FP16-shifted storage within
the stated velocity tolerances of the FP64 reference.
"""
import numpy as np
import pytest

from test_oracle_walls import GUO_CASES, guo_case

pytestmark = pytest.mark.gpu


def make_wall_lattice(c, storage="f64"):
    from biblbm_amd.lattice import Lattice
    from biblbm_amd.walls import Boundary, GuoWall
    nx, ny, nz = c["f0"].shape[:3]
    L = Lattice(nx, ny, nz, periodic=[bool(p) for p in c["periodic"]], storage=storage, omega=c["omega"])
    L.set_mask(c["mask"])
    L.set_body_force(c["force"][0, 0, 0])
    L.put_populations(c["f0"])
    ob = c["boundary"]
    b = Boundary(ob.kind, type(Boundary().p)(*ob.p))
    W = GuoWall(L, b)
    W.set_node_velocities(c["wall_vel"])
    W.insert()
    return L, W


def check_cells(c):
    shape = c["f0"].shape[:3]
    wall = np.zeros(shape, bool)
    wall[np.unravel_index(c["cell"], shape)] = True
    return (c["mask"] == 0) | wall


@pytest.mark.parametrize("name", GUO_CASES)
def test_guo_wall_nodes_match_reference(device, name):
    c = guo_case(name)
    L, W = make_wall_lattice(c)
    n = W.nodes()
    assert np.array_equal(n["cell"], c["cell"])
    assert np.array_equal(n["dir"], c["dir"])
    assert np.array_equal(n["delta"], c["delta"])
    assert np.array_equal(n["wall_pos"], c["wall_pos"])
    W.close()
    L.close()


@pytest.mark.parametrize("name", GUO_CASES)
def test_guo_wall_f64_matches_reference(device, name):
    c = guo_case(name)
    L, W = make_wall_lattice(c)
    chk = check_cells(c)
    L.collide_and_stream(1)
    s1 = np.abs(c["f1"][chk]).max()
    e1 = np.abs(L.get_populations() - c["f1"])[chk].max() / s1
    L.collide_and_stream(c["steps"] - 1)
    sN = np.abs(c["fN"][chk]).max()
    eN = np.abs(L.get_populations() - c["fN"])[chk].max() / sN
    print(name, "rel err after 1 step", e1, "after", c["steps"], eN)
    assert e1 <= 1e-12 and eN <= 1e-12
    W.close()
    L.close()


@pytest.mark.parametrize("storage,tol", [("f32", 2e-5), ("f16s", 5e-3)])
@pytest.mark.parametrize("name", GUO_CASES)
def test_guo_wall_reduced_storage(device, storage, tol, name):
    """Velocity after N steps against the reference, relative to max |u|."""
    from oracle import oracle_py as O
    c = guo_case(name)
    L, W = make_wall_lattice(c, storage)
    L.collide_and_stream(c["steps"])
    rho, u = L.moments()
    _, uref = O.lbm_moments(c["fN"], c["force"], c["mask"])
    fluid = c["mask"] == 0
    err = np.abs(u - uref)[fluid].max() / np.abs(uref[fluid]).max()
    print(storage, name, "max rel velocity error", err)
    assert err <= tol
    W.close()
    L.close()


def test_guo_wall_runs_in_initialize_and_survives_lattice_destruction(device):
    """bl_lattice_initialize runs the wall pass once (as lattice.initialize()
    runs the integrated processors); a wall whose lattice is gone reports an
    error instead of touching freed memory."""
    from biblbm_amd.capi import BlError, call
    from oracle import oracle_py as O
    c = guo_case("pipe_moving")
    L, W = make_wall_lattice(c)
    call("bl_lattice_initialize", L.handle)
    w = O.guo_wall_nodes(c["f0"].shape[:3], c["boundary"])
    ref = O.guo_wall_apply(c["f0"], c["periodic"], w, c["wall_vel"], c["mask"])
    chk = check_cells(c)
    assert np.abs(L.get_populations() - ref)[chk].max() <= 1e-12 * np.abs(ref[chk]).max()
    L.close()
    with pytest.raises(BlError):
        W.set_wall_velocity([0.0, 0.0, 0.0])
    W.close()


def test_guo_wall_rejects_a_lattice_too_thin_for_the_extrapolation(device):
    """A node whose second neighbour along its link leaves a non-periodic
    lattice is refused at creation."""
    from biblbm_amd.capi import BlError
    from biblbm_amd.lattice import Lattice
    from biblbm_amd.walls import Boundary, GuoWall
    L = Lattice(4, 3, 4, periodic=(True, False, True), omega=1.0)
    with pytest.raises(BlError):
        GuoWall(L, Boundary.plates(1.5, 2.5))
    L.close()


def test_guo_wall_nodes_on_the_surface_keep_the_sequential_order(device):
    """Integer centre and radius put lattice nodes exactly on the pipe
    surface: some wall nodes then read other wall nodes.  The pass keeps the
    reference's node-by-node order (the oracle runs it in place, in node
    order); populations after several steps agree to round-off."""
    from biblbm_amd.lattice import Lattice
    from biblbm_amd.walls import Boundary, GuoWall, outside_mask
    from oracle import oracle_py as O
    shape, per = (6, 21, 21), (1, 0, 0)
    b, ob = Boundary.pipe(10.0, 10.0, 8.0), O.Boundary.pipe(10.0, 10.0, 8.0)
    mask = outside_mask(b, shape, 1.0)
    rng = np.random.default_rng(12)
    u = rng.uniform(-0.02, 0.02, shape + (3,))
    f0 = O.lbm_init_equilibrium(rng.uniform(0.99, 1.01, shape), u)
    w = O.guo_wall_nodes(shape, ob)
    c = np.unravel_index(w["cell"], shape)
    node = np.zeros(shape, bool)
    node[c] = True
    reads_node = False
    C = np.array([[0, 0, 0], [-1, 0, 0], [0, -1, 0], [0, 0, -1], [-1, -1, 0], [-1, 1, 0], [-1, 0, -1],
                            [-1, 0, 1], [0, -1, -1], [0, -1, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0],
                            [1, -1, 0], [1, 0, 1], [1, 0, -1], [0, 1, 1], [0, 1, -1]])
    for k in range(len(w["cell"])):
        p = np.array([c[0][k], c[1][k], c[2][k]])
        for m in (1, 2):
            q = p + m * C[w["dir"][k]]
            q[0] %= shape[0]
            reads_node = reads_node or bool(node[tuple(q)])
    assert reads_node
    vel = np.tile([0.01, 0.0, 0.0], (len(w["cell"]), 1))
    L = Lattice(*shape, periodic=per, omega=1 / 0.8)
    L.set_mask(mask)
    L.put_populations(f0)
    W = GuoWall(L, b)
    W.set_wall_velocity([0.01, 0.0, 0.0])
    W.insert()
    steps = 6
    L.collide_and_stream(steps)
    ref = O.lbm_run(f0, np.zeros(shape + (3,)), np.zeros(shape), mask, per, 1 / 0.8, 1 / 0.8, steps, guo=(w, vel))
    chk = (mask == 0) | node
    err = np.abs(L.get_populations() - ref)[chk].max() / np.abs(ref[chk]).max()
    print("surface nodes: rel err", err)
    assert err <= 1e-12
    W.close()
    L.close()


# ------------------------------------------------------------------ near-wall immersed boundary
def _capsule(center_y, radius=3.5, level=2, x=10.2, z=9.7):
    from biblbm_amd import meshes
    rest, tri = meshes.icosphere(level, radius)
    pos = rest + [x, center_y, z]
    return rest, tri, pos


@pytest.mark.parametrize("kernel", ["peskin", "roma"])
def test_near_wall_coupled_steps_plates(device, kernel):
    """A Skalak capsule next to a bounce-back plate with the IBM boundary set
    (plates at the bounce-back half-way planes): every near-wall branch
    (kernel only, blend, top hat only) is taken; GPU trajectory and flow
    against the oracle's near-wall coupled steps."""
    from biblbm_amd.lattice import Lattice
    from biblbm_amd.ibm import ImmersedBoundary
    from biblbm_amd.walls import Boundary
    from oracle import oracle_py as O
    shape, per = (20, 18, 20), (1, 0, 1)
    rest, tri, pos = _capsule(0.5 + 3.5 + 0.4)
    mask = np.zeros(shape, np.uint8)
    mask[:, 0, :] = mask[:, -1, :] = 1
    rng = np.random.default_rng(4)
    f0 = O.lbm_init_equilibrium(rng.uniform(0.995, 1.005, shape), rng.uniform(-0.01, 0.01, shape + (3,)))
    body = (3e-6, 0.0, 0.0)
    L = Lattice(*shape, periodic=per, omega=1 / 0.9)
    L.set_mask(mask)
    L.set_body_force(body)
    L.put_populations(f0)
    ib = ImmersedBoundary(L, kernel)
    sid = ib.add_shape(rest, tri)
    ib.add_particle(sid, 0, [0.02, 10.0], pos)
    ib.init()
    ib.set_boundary(Boundary.plates(0.5, 16.5))
    steps = 5
    ib.coupled_steps(steps, use_graph=False)
    parts = [O.Particle(rest, tri, 0, [0.02, 10.0], pos)]
    ob = O.Boundary.plates(0.5, 16.5)
    f, force = f0, np.broadcast_to(np.asarray(body), shape + (3,)).copy()
    for _ in range(steps):
        f, force, _ = O.coupled_step(f, force, np.zeros(shape), mask, per, 1 / 0.9, 1 / 0.9, body, parts, kernel,
                                     boundary=ob)
    ref = parts[0].pos
    err = np.linalg.norm(ib.positions() - ref) / np.linalg.norm(ref - pos)
    _, u = L.moments()
    _, uref = O.lbm_moments(f, force, mask)
    err_u = np.linalg.norm(u - uref) / np.linalg.norm(uref)
    print(kernel, "position error / displacement", err, "flow rel L2", err_u)
    assert err <= 1e-9 and err_u <= 1e-11
    # the capsule really straddles all three branches
    hs = {"peskin": 2.0, "roma": 1.5}[kernel]
    w = pos[:, 1] - 0.5 - hs
    assert (w <= 0).any() and ((w > 0) & (w < 1)).any() and (w >= 1).any()


def test_near_wall_vertex_reflection(device):
    """A vertex 0.005 above the lower plate in a flow towards it crosses the
    plate within one step and is reflected back, on the GPU as in the
    oracle."""
    from biblbm_amd.lattice import Lattice
    from biblbm_amd.ibm import ImmersedBoundary
    from biblbm_amd.walls import Boundary
    from oracle import oracle_py as O
    shape, per = (16, 14, 16), (1, 0, 1)
    rest, tri, _ = _capsule(0.0, radius=3.0, x=0.0, z=0.0)
    pos = rest + [8.1, 0.505 - rest[:, 1].min(), 7.9]
    mask = np.zeros(shape, np.uint8)
    mask[:, 0, :] = mask[:, -1, :] = 1
    u0 = np.zeros(shape + (3,))
    u0[..., 1] = -0.01
    f0 = O.lbm_init_equilibrium(np.ones(shape), u0)
    L = Lattice(*shape, periodic=per, omega=1.0)
    L.set_mask(mask)
    L.put_populations(f0)
    ib = ImmersedBoundary(L, "peskin")
    sid = ib.add_shape(rest, tri)
    ib.add_particle(sid, 0, [0.02, 10.0], pos)
    ib.init()
    ib.set_boundary(Boundary.plates(0.5, 12.5))
    ib.coupled_steps(1, use_graph=False)
    parts = [O.Particle(rest, tri, 0, [0.02, 10.0], pos)]
    f, force, vels = O.coupled_step(f0, np.zeros(shape + (3,)), np.zeros(shape), mask, per, 1.0, 1.0,
                                    (0, 0, 0), parts, "peskin", boundary=O.Boundary.plates(0.5, 12.5))
    crossed = (pos + vels[0])[:, 1] < 0.5
    assert crossed.any()
    got = ib.positions()
    assert np.abs(got - parts[0].pos).max() <= 1e-12
    assert got[:, 1].min() > 0.5


def test_pipe_guo_wall_with_near_wall_capsule(device):
    """rbc_in_pipe's second-order set-up in miniature: Guo curved wall on the
    pipe, no dynamics beyond it, a capsule near the wall with the pipe as
    the IBM boundary; four coupled steps against the oracle."""
    from biblbm_amd.capi import call
    from biblbm_amd.lattice import Lattice
    from biblbm_amd.ibm import ImmersedBoundary
    from biblbm_amd.walls import Boundary, GuoWall, outside_mask
    from oracle import oracle_py as O
    shape, per = (12, 24, 24), (1, 0, 0)
    b = Boundary.pipe(11.5, 11.3, 10.0, 12.0)
    ob = O.Boundary.pipe(11.5, 11.3, 10.0, 12.0)
    mask = outside_mask(b, shape, 1.0)
    body = (2e-5, 0.0, 0.0)
    L = Lattice(*shape, periodic=per, omega=1 / 0.85)
    L.set_mask(mask)
    L.set_body_force(body)
    f0 = O.lbm_init_equilibrium(np.ones(shape), np.zeros(shape + (3,)))
    L.put_populations(f0)
    W = GuoWall(L, b)
    W.insert()
    call("bl_lattice_initialize", L.handle)
    w = O.guo_wall_nodes(shape, ob)
    assert np.array_equal(W.nodes()["cell"], w["cell"])
    vel0 = np.zeros((len(w["cell"]), 3))
    f = O.guo_wall_apply(f0, per, w, vel0, mask)
    rest, tri, _ = _capsule(0.0, radius=3.0, x=0.0, z=0.0)
    pos = rest + [6.3, 11.5 + 5.6, 11.3]
    ib = ImmersedBoundary(L, "peskin")
    sid = ib.add_shape(rest, tri)
    ib.add_particle(sid, 1, [0.01, 10.0], pos)
    ib.init()
    ib.set_boundary(b)
    steps = 4
    ib.coupled_steps(steps, use_graph=True)
    parts = [O.Particle(rest, tri, 1, [0.01, 10.0], pos)]
    force = np.broadcast_to(np.asarray(body), shape + (3,)).copy()
    for _ in range(steps):
        f, force, _ = O.coupled_step(f, force, np.zeros(shape), mask, per, 1 / 0.85, 1 / 0.85, body, parts,
                                     "peskin", boundary=ob, guo=(w, vel0))
    ref = parts[0].pos
    err = np.linalg.norm(ib.positions() - ref) / np.linalg.norm(ref - pos)
    fluid = mask == 0
    err_f = np.abs(L.get_populations() - f)[fluid].max() / np.abs(f[fluid]).max()
    print("pipe: position error / displacement", err, "population rel err", err_f)
    assert err <= 1e-9 and err_f <= 1e-11
    W.close()


def test_second_order_pipe_workload_matches_oracle(device):
    """The benchmark's pipe with the reference prog's default wall (Guo
    curved wall, no dynamics beyond it) at a reduced size: three coupled
    steps of the red cell (Skalak + bending + area + volume) against the
    oracle started from the same state."""
    from biblbm_amd import workloads
    from oracle import oracle_py as O
    W = workloads.rbc_in_pipe("f64", R=8, length=20, level=2, cell_radius=3.5, second_order=True)
    L, ib = W.lattice, W.ibm
    f0 = L.get_populations()
    pos0 = ib.positions()
    p = W.params
    c = (2 * 8 + 3 - 1) / 2.0
    ob = O.Boundary.pipe(c, c, 8.0, 20.0)
    shape = L.shape
    w = O.guo_wall_nodes(shape, ob)
    mask = np.zeros(shape, np.uint8)
    from biblbm_amd.walls import Boundary, outside_mask
    mask = outside_mask(Boundary.pipe(c, c, 8.0, 20.0), shape, 1.0)
    from biblbm_amd import meshes
    v, t = meshes.biconcave(2, 3.5)
    vol = meshes.signed_volume(v, t)
    params = workloads.skalak_rbc_params(p["G"], 10.0, 3.5, vol, p["k_bend"])
    parts = [O.Particle(v, t, 3, params, pos0)]
    steps = 3
    ib.coupled_steps(steps, use_graph=False)
    f = f0
    body = (p["g"], 0.0, 0.0)
    force = np.broadcast_to(np.asarray(body), shape + (3,)).copy()
    for _ in range(steps):
        f, force, _ = O.coupled_step(f, force, np.zeros(shape), mask, (1, 0, 0), 1 / 0.9, 1 / 0.9, body, parts,
                                     "peskin", guo=(w, np.zeros((len(w["cell"]), 3))))
    ref = parts[0].pos
    err = np.linalg.norm(ib.positions() - ref) / np.linalg.norm(ref - pos0)
    fluid = mask == 0
    err_f = np.abs(L.get_populations() - f)[fluid].max() / np.abs(f[fluid]).max()
    print("second-order pipe: position error / displacement", err, "population rel err", err_f)
    assert err <= 1e-9 and err_f <= 1e-11
    ib.close()
    W.guo.close()
    L.close()
