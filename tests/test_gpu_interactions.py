"""Interaction forces on the GPU (libbiblbm.so through the C-ABI) against the
reference's golden vectors and the FP64 oracle.

Bar: FP64 collision-handler forces within 1e-13 of the reference's
(relative to their largest magnitude), on exactly the reference's set of
interacting nodes; the same against the oracle for a 25 680-vertex
suspension of capsules, with the pair forces summing to zero and two
evaluations identical bit for bit; coupled steps with a wall-interaction
decorator and a particle-particle force within the coupled-step bar of the
near-wall tests (positions 1e-9 of the displacement, flow 1e-11 relative
L2), eagerly and from the captured graph.
"""
import numpy as np
import pytest

from test_oracle_interactions import PP_CASES, pp_case

pytestmark = pytest.mark.gpu

FAR = np.array([[-50.0, -50.0, -50.0], [-50.0, -49.0, -50.0]])


def _fan(n):
    return np.array([[0, i, i + 1] for i in range(1, n - 1)], np.int32)


def _golden_particles(ib, c):
    """One particle per owner of the golden's nodes, in node order: a fan
    mesh over the owner's nodes; single nodes get two companions far outside
    the collision box.  Returns the device positions and the device index of
    every golden node."""
    pos, obj = c["pos"], c["obj"]
    dev_pos, index = [], []
    start = 0
    while start < len(pos):
        end = start
        while end < len(pos) and obj[end] == obj[start]:
            end += 1
        p = pos[start:end]
        index.extend(range(sum(len(q) for q in dev_pos), sum(len(q) for q in dev_pos) + len(p)))
        if len(p) >= 3:
            rest, tri, cur = p, _fan(len(p)), p
        else:
            assert len(p) == 1
            rest = np.array([p[0], p[0] + [1, 0, 0], p[0] + [0, 1, 0]])
            tri = np.array([[0, 1, 2]], np.int32)
            cur = np.vstack([p, FAR])
        sid = ib.add_shape(rest, tri)
        ib.add_particle(sid, 0, [0.02, 10.0], cur)
        dev_pos.append(cur)
        start = end
    return np.vstack(dev_pos), np.array(index)


@pytest.mark.parametrize("name", PP_CASES)
def test_collision_forces_match_reference(device, name):
    from biblbm_amd.lattice import Lattice
    from biblbm_amd.ibm import ImmersedBoundary
    from biblbm_amd.potentials import Potential
    from biblbm_amd.walls import Boundary
    c = pp_case(name)
    L = Lattice(*c["shape"], periodic=tuple(c["periodic"]), omega=1.0)
    ib = ImmersedBoundary(L, "peskin")
    allpos, idx = _golden_particles(ib, c)
    ib.init()
    ib.set_positions(allpos)
    u = Potential(c["u"].kind, c["u"].p)
    wall = None if c["wall"] is None else Boundary(c["wall"].kind, c["wall"].p)
    ib.compute_collision_forces(u, wall)
    f = ib.forces()[idx]
    ref = c["force"]
    scale = np.abs(ref).max()
    err = np.abs(f - ref).max() / scale
    print(name, "nodes", len(ref), "interacting", np.count_nonzero(np.abs(ref).sum(1)), "rel err", err)
    assert np.array_equal(np.abs(f).sum(1) != 0, np.abs(ref).sum(1) != 0)
    assert err <= 1e-13
    ib.close()
    L.close()


def test_pp_force_cutoff_limit(device):
    from biblbm_amd.capi import BlError
    from biblbm_amd.lattice import Lattice
    from biblbm_amd.ibm import ImmersedBoundary
    from biblbm_amd.potentials import MorsePotential, SpringPotential
    L = Lattice(8, 8, 8, omega=1.0)
    ib = ImmersedBoundary(L, "roma")  # half support 1.5: cut-offs up to 1.0
    ib.add_pp_force(SpringPotential(1.0, 0.1))
    with pytest.raises(BlError, match="half support"):
        ib.add_pp_force(MorsePotential(1.2, 1.0, 1e-3))
    with pytest.raises(BlError, match="positive"):
        ib.add_pp_force(SpringPotential(-1.0, 0.1))
    ib.close()
    L.close()


def _suspension(n_caps=40, box=48, seed=3):
    from biblbm_amd import meshes
    rest, tri = meshes.icosphere(3, 4.0)
    rng = np.random.default_rng(seed)
    pos = []
    for k in range(n_caps):
        ctr = rng.uniform(0, box, 3)
        A = np.eye(3) + rng.uniform(-0.08, 0.08, (3, 3))
        pos.append(rest @ A.T + ctr)
    return rest, tri, pos


def test_suspension_collision_forces_match_oracle(device):
    from biblbm_amd.lattice import Lattice
    from biblbm_amd.ibm import ImmersedBoundary
    from biblbm_amd.potentials import MorsePotential
    from oracle import oracle_py as O
    box = 48
    rest, tri, pos = _suspension(box=box)
    L = Lattice(box, box, box, periodic=(1, 1, 1), omega=1.0)
    ib = ImmersedBoundary(L, "peskin")
    sid = ib.add_shape(rest, tri)
    for p in pos:
        ib.add_particle(sid, 0, [0.02, 10.0], p)
    ib.init()
    allpos = np.vstack(pos)
    ib.set_positions(allpos)
    u = MorsePotential(1.5, 1.0, 1e-3)
    ib.compute_collision_forces(u)
    f1 = ib.forces()
    ib.compute_collision_forces(u)
    f2 = ib.forces()
    obj = np.repeat(np.arange(len(pos), dtype=np.int32), len(rest))
    ref, pairs = O.collision_forces((box,) * 3, [1, 1, 1], O.Potential.morse(1.5, 1.0, 1e-3), allpos, obj)
    scale = np.abs(ref).max()
    err = np.abs(f1 - ref).max() / scale
    print("vertices", len(allpos), "pairs", pairs, "interacting", np.count_nonzero(np.abs(ref).sum(1)),
          "rel err", err)
    assert pairs > 1000
    assert np.array_equal(f1, f2)
    assert err <= 1e-13
    assert np.abs(f1.sum(0)).max() <= 1e-12 * scale
    ib.close()
    L.close()


@pytest.mark.parametrize("use_graph", [False, True])
def test_coupled_steps_with_wall_and_pp_forces(device, use_graph):
    """Two Skalak capsules 0.25 apart, resting 1.0 above a bounce-back plate:
    the Morse wall decorator and the Morse particle-particle force both act
    from the first step; GPU trajectory and flow against the oracle."""
    from biblbm_amd import meshes
    from biblbm_amd.lattice import Lattice
    from biblbm_amd.ibm import ImmersedBoundary
    from biblbm_amd.potentials import MorsePotential
    from biblbm_amd.walls import Boundary
    from oracle import oracle_py as O
    shape, per = (24, 20, 18), (1, 0, 1)
    rest, tri = meshes.icosphere(2, 3.5)
    y = 0.5 + 3.5 + 1.0
    pos = [rest + [7.0, y, 9.0], rest + [7.0 + 7.0 + 0.25, y + 0.3, 9.2]]
    mask = np.zeros(shape, np.uint8)
    mask[:, 0, :] = mask[:, -1, :] = 1
    rng = np.random.default_rng(7)
    f0 = O.lbm_init_equilibrium(rng.uniform(0.995, 1.005, shape), rng.uniform(-0.005, 0.005, shape + (3,)))
    body = (2e-6, 0.0, 0.0)
    L = Lattice(*shape, periodic=per, omega=1 / 0.9)
    L.set_mask(mask)
    L.set_body_force(body)
    L.put_populations(f0)
    ib = ImmersedBoundary(L, "peskin")
    sid = ib.add_shape(rest, tri)
    for p in pos:
        ib.add_particle(sid, 0, [0.02, 10.0], p)
    ib.init()
    ib.add_wall_interaction(Boundary.plates(0.5, 18.5), MorsePotential(2.0, 2.0, 1e-4))
    ib.add_pp_force(MorsePotential(1.0, 1.0, 1e-3))
    steps = 6
    ib.coupled_steps(steps, use_graph=use_graph)
    parts = [O.Particle(rest, tri, 0, [0.02, 10.0], p) for p in pos]
    wi = [(O.Boundary.plates(0.5, 18.5), O.Potential.morse(2.0, 2.0, 1e-4))]
    pp = [O.Potential.morse(1.0, 1.0, 1e-3)]
    f, force = f0, np.broadcast_to(np.asarray(body), shape + (3,)).copy()
    F0 = O.vertex_forces(parts, shape, per, wi, pp)
    for _ in range(steps):
        f, force, _ = O.coupled_step(f, force, np.zeros(shape), mask, per, 1 / 0.9, 1 / 0.9, body, parts, "peskin",
                                     wall_interactions=wi, pp_forces=pp)
    ref = np.vstack([p.pos for p in parts])
    p0 = np.vstack(pos)
    err = np.linalg.norm(ib.positions() - ref) / np.linalg.norm(ref - p0)
    _, uu = L.moments()
    _, uref = O.lbm_moments(f, force, mask)
    err_u = np.linalg.norm(uu - uref) / np.linalg.norm(uref)
    print("graph" if use_graph else "eager", "position error / displacement", err, "flow rel L2", err_u)
    assert err <= 1e-9 and err_u <= 1e-11
    # both interactions really act at the start
    fresh = [O.Particle(rest, tri, 0, [0.02, 10.0], p) for p in pos]
    d_wall_pp = np.vstack(F0) - np.vstack(O.vertex_forces(fresh, shape, per))
    assert np.count_nonzero(np.abs(d_wall_pp).sum(1) > 1e-9) > 10
    ib.close()
    L.close()


def test_compute_forces_stage_with_interactions(device):
    """The force stage alone (decorators, membrane, particle-particle) on a
    suspension of touching capsules near plates, against the oracle's
    compute_forces."""
    from biblbm_amd.lattice import Lattice
    from biblbm_amd.ibm import ImmersedBoundary
    from biblbm_amd.potentials import MorsePotential, SpringPotential
    from biblbm_amd.walls import Boundary
    from oracle import oracle_py as O
    box = 48
    rest, tri, pos = _suspension(n_caps=30, box=box, seed=5)
    pos = [np.clip(p, 0.5, box - 1.5) for p in pos]  # some vertices pressed onto the plates
    L = Lattice(box, box, box, periodic=(1, 0, 1), omega=1.0)
    ib = ImmersedBoundary(L, "peskin")
    sid = ib.add_shape(rest, tri)
    for p in pos:
        ib.add_particle(sid, 0, [0.02, 10.0], p)
    ib.init()
    ib.set_positions(np.vstack(pos))
    ib.add_wall_interaction(Boundary.plates(0.5, box - 1.5), MorsePotential(2.0, 2.0, 1e-4))
    ib.add_pp_force(MorsePotential(1.0, 1.0, 1e-3))
    ib.add_pp_force(SpringPotential(0.8, 0.05))
    ib.compute_forces()
    f = ib.forces()
    parts = [O.Particle(rest, tri, 0, [0.02, 10.0], p) for p in pos]
    wi = [(O.Boundary.plates(0.5, box - 1.5), O.Potential.morse(2.0, 2.0, 1e-4))]
    pp = [O.Potential.morse(1.0, 1.0, 1e-3), O.Potential.spring(0.8, 0.05)]
    ref = np.vstack(O.vertex_forces(parts, (box, box, box), (1, 0, 1), wi, pp))
    bare = np.vstack(O.vertex_forces(parts, (box, box, box), (1, 0, 1)))
    err = np.abs(f - ref).max() / np.abs(ref).max()
    print("rel err", err, "vertices changed by the interactions", np.count_nonzero(np.abs(ref - bare).sum(1) > 0))
    assert np.count_nonzero(np.abs(ref - bare).sum(1) > 0) > 100
    assert err <= 1e-12
    ib.close()
    L.close()
