"""Rigid particles and the rigid-growth packing on the device, through the
C-ABI, against the reference (oracle/ref/ref_rigid.cpp: the growth loop of
progs/common/create_particles.h:129-155 run with the reference's own
RigidParticle3D, CollisionHandler and SpringPotential) and the oracle.

- bl_rigid_grow over the whole 401-iteration loop (14 particles of a
  314-vertex sphere growing from scale 0.2 to 1 in a pipe periodic along x):
  centres, velocities, angular velocities, quaternions and vertices at the
  end; the first 11 iterations step by step (bl_rigid_set_scale,
  bl_ibm_compute_collision_forces, bl_rigid_move) against the reference's
  state after iteration 10.  Tolerances: the device's collision forces fuse
  a*b+c where the reference's do not (last-bit differences, grown by 401
  iterations of contacts).
- Each move against the oracle's rigid_move on the same state, positions
  and device forces (FP64, 1e-12).
- The rest volume and a centre that wraps across the periodic face; a
  boundary holding rigid particles refuses the fluid-coupling stages.
"""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def _setup(name="pipe14"):
    from biblbm_amd.ibm import ImmersedBoundary
    from biblbm_amd.lattice import Lattice
    from biblbm_amd.potentials import SpringPotential
    from biblbm_amd.walls import Boundary
    g = lambda n: load_golden("rigid", name, n + ".npy")
    P = g("params")
    shape = tuple(int(x) for x in P[0:3])
    L = Lattice(*shape, periodic=tuple(int(x) for x in P[3:6]), storage="f64", omega=1.0)
    ib = ImmersedBoundary(L, "peskin")
    sid = ib.add_shape(g("rest"), g("tri"))
    for k in range(int(P[17])):
        ib.add_rigid(sid, g("com0")[k], g("quat0")[k], P[12], P[15], P[16])
    ib.init()
    u = SpringPotential(P[10], P[11])
    wall = Boundary.pipe(P[6], P[7], P[8], P[9])
    return dict(g=g, P=P, L=L, ib=ib, sid=sid, u=u, wall=wall, n=int(P[17]))


def _states(ib, n):
    s = [ib.rigid_state(k) for k in range(n)]
    return {k: np.array([d[k] for d in s]) for k in ("com", "vel", "angvel", "quat", "scale")}


def _close(c):
    c["ib"].close()
    c["L"].close()


def test_rest_volume_matches_reference(device):
    c = _setup()
    assert c["ib"].shape_info(c["sid"])["volume"] == pytest.approx(c["P"][18], rel=1e-14)
    _close(c)


def test_first_iterations_match_reference(device):
    c = _setup()
    ib, P = c["ib"], c["P"]
    n_iter, init_s, every = int(P[13]), P[12], int(P[14])
    for it in range(11):
        if it % every == 0:
            ib.rigid_set_scale(init_s + (1 - init_s) * (it / n_iter))
        ib.compute_collision_forces(c["u"], c["wall"])
        ib.rigid_move()
    s, g = _states(ib, c["n"]), c["g"]
    np.testing.assert_allclose(s["com"], g("com_10"), rtol=0, atol=1e-12)
    np.testing.assert_allclose(s["quat"], g("quat_10"), rtol=0, atol=1e-13)
    np.testing.assert_array_equal(s["scale"], g("scale_10"))
    np.testing.assert_allclose(ib.positions(), g("pos_10"), rtol=0, atol=1e-12)
    _close(c)


def test_growth_loop_matches_reference(device):
    c = _setup()
    ib, P, g = c["ib"], c["P"], c["g"]
    ib.rigid_grow(int(P[13]), P[12], c["u"], c["wall"], every=int(P[14]))
    s = _states(ib, c["n"])
    np.testing.assert_array_equal(s["scale"], g("scale_end"))
    np.testing.assert_allclose(s["com"], g("com_end"), rtol=0, atol=1e-8)
    np.testing.assert_allclose(s["vel"], g("vel_end"), rtol=0, atol=1e-10)
    np.testing.assert_allclose(s["angvel"], g("angvel_end"), rtol=0, atol=1e-10)
    np.testing.assert_allclose(s["quat"], g("quat_end"), rtol=0, atol=1e-8)
    np.testing.assert_allclose(ib.positions(), g("pos_end"), rtol=0, atol=1e-8)
    _close(c)


def test_moves_match_oracle(device):
    from oracle import oracle_py as O
    c = _setup()
    ib, g, P = c["ib"], c["g"], c["P"]
    rest, inertia = g("rest"), g("inertia")
    nv = len(rest)
    ib.rigid_set_scale(0.9)
    for step in range(3):
        ib.compute_collision_forces(c["u"], c["wall"])
        before = _states(ib, c["n"])
        pos0, f = ib.positions(), ib.forces()
        ib.rigid_move()
        after, pos1 = _states(ib, c["n"]), ib.positions()
        for k in range(c["n"]):
            st = np.concatenate([before["com"][k], before["vel"][k], before["angvel"][k], before["quat"][k],
                                 [before["scale"][k], P[15], P[16]]])
            sl = slice(k * nv, (k + 1) * nv)
            st1, p1 = O.rigid_move(rest, inertia, P[18], [1, 0, 0], c["L"].shape, st, pos0[sl], f[sl])
            np.testing.assert_allclose(after["com"][k], st1[0:3], rtol=0, atol=1e-12)
            np.testing.assert_allclose(after["vel"][k], st1[3:6], rtol=1e-12, atol=1e-18)
            np.testing.assert_allclose(after["angvel"][k], st1[6:9], rtol=1e-12, atol=1e-18)
            np.testing.assert_allclose(after["quat"][k], st1[9:13], rtol=0, atol=1e-14)
            np.testing.assert_allclose(pos1[sl], p1, rtol=0, atol=1e-12)
    assert np.abs(after["vel"]).max() > 0
    _close(c)


def test_centre_wraps_across_the_periodic_face(device):
    from biblbm_amd.ibm import ImmersedBoundary
    from biblbm_amd.lattice import Lattice
    from biblbm_amd import meshes
    v, t = meshes.icosphere(2, 2.0)
    L = Lattice(20, 12, 12, periodic=(1, 0, 0), storage="f64", omega=1.0)
    ib = ImmersedBoundary(L, "peskin")
    sid = ib.add_shape(v, t)
    ib.add_rigid(sid, (20.3, 6.0, 6.0))  # past the face: init maps it
    ib.init()
    s = ib.rigid_state(0)
    assert s["com"][0] == pytest.approx(0.3, abs=1e-12)
    np.testing.assert_allclose(ib.positions(), v + (0.3, 6.0, 6.0), rtol=0, atol=1e-12)
    from biblbm_amd import capi
    with pytest.raises(capi.BlError):
        ib.coupled_steps(1, use_graph=False)
    with pytest.raises(capi.BlError):
        ib.step()
    ib.close()
    L.close()
