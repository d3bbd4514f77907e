"""The oracle's rigid bodies (growth packing) against the reference and
against closed forms.

- Reference: the growth loop of progs/common/create_particles.h:129-155 run
  with the reference's own RigidParticle3D, CollisionHandler, PipeBoundary
  and SpringPotential (oracle/ref/ref_rigid.cpp: 14 particles of a 314-vertex
  sphere in a pipe periodic along x, 401 iterations from scale 0.2 to 1):
  centres, velocities, angular velocities, quaternions and vertex positions
  after 10 iterations and at the end, bit for bit, and the rest-shape
  inertia.
- Closed forms: a body at rest stays at rest on com + R (s v); a uniform
  force with damping d gives v_k = F / (m d) (1 - (1 - d)^k); a body with
  isotropic inertia and no torque spins at constant rate (quaternion after k
  steps = rotation by k |w| about w); the centre of mass wraps onto
  [-1/2, n - 1/2) of a periodic axis with the vertices.
"""
import numpy as np
import pytest

from conftest import load_golden
from oracle import oracle_py as O


def golden(name):
    g = lambda n: load_golden("rigid", name, n + ".npy")
    P = g("params")
    n = int(P[17])
    states = np.array([O.rigid_state(g("com0")[k], g("quat0")[k], P[12], P[15], P[16]) for k in range(n)])
    return dict(g=g, P=P, n=n, rest=g("rest"), tri=g("tri"), inertia=g("inertia"), volume=P[18],
                shape=tuple(int(x) for x in P[0:3]), periodic=[int(x) for x in P[3:6]],
                boundary=O.Boundary.pipe(P[6], P[7], P[8], P[9]), u=O.Potential.spring(P[10], P[11]),
                initial_scale=P[12], n_iter=int(P[13]), every=int(P[14]), states=states)


def test_shape_inertia_matches_reference():
    c = golden("pipe14")
    np.testing.assert_allclose(O.shape_inertia(c["rest"], c["tri"]), c["inertia"], rtol=1e-14, atol=0)


def test_growth_loop_matches_reference():
    c = golden("pipe14")
    states, pos, rec = O.rigid_grow(c["rest"], c["inertia"], c["volume"], c["shape"], c["periodic"], c["states"],
                                    c["u"], c["boundary"], c["n_iter"], c["initial_scale"], c["every"], record=(10,))
    g = c["g"]
    for tag, (st, ps) in (("10", rec[10]), ("end", (states, pos))):
        # bit for bit: same operations in the same order, same libm
        np.testing.assert_array_equal(st[:, 0:3], g("com_" + tag))
        np.testing.assert_array_equal(st[:, 3:6], g("vel_" + tag))
        np.testing.assert_array_equal(st[:, 6:9], g("angvel_" + tag))
        np.testing.assert_array_equal(st[:, 9:13], g("quat_" + tag))
        np.testing.assert_array_equal(st[:, 13], g("scale_" + tag))
        np.testing.assert_array_equal(ps, g("pos_" + tag))
    # the particles did collide: they moved and turned
    assert np.abs(g("vel_end")).max() > 1e-3 and np.abs(g("angvel_end")).max() > 1e-4


def _sphere():
    from biblbm_amd import meshes
    v, t = meshes.icosphere(2, 3.0)
    return v, t


def test_body_at_rest_stays_on_its_rest_shape():
    v, t = _sphere()
    q = np.array([np.cos(0.4), 0.6 * np.sin(0.4), 0.0, 0.8 * np.sin(0.4)])
    st = O.rigid_state((5.0, 6.0, 7.0), q, 0.7, 1.0, 0.1)
    pos = O.rigid_update(v, st)
    st2, pos2 = O.rigid_move(v, O.shape_inertia(v, t), 100.0, [0, 0, 0], (20, 20, 20), st, pos, np.zeros_like(pos))
    np.testing.assert_array_equal(st2, st)
    np.testing.assert_array_equal(pos2, pos)
    # com + R (s v) with R the rotation of the quaternion
    w, x, y, z = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (w * y + x * z)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (w * x + y * z), 1 - 2 * (x * x + y * y)]])
    np.testing.assert_allclose(pos, (5.0, 6.0, 7.0) + 0.7 * v @ R.T, rtol=0, atol=1e-13)
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-15)


def test_uniform_force_with_damping():
    v, t = _sphere()
    v = v - v.mean(axis=0)  # no torque from a uniform force
    F = np.array([3e-3, -1e-3, 2e-3])
    vol, d, s, rho = 50.0, 0.1, 0.5, 2.0
    m = s ** 3 * vol * rho
    st = O.rigid_state((10.0, 10.0, 10.0), (1, 0, 0, 0), s, rho, d)
    pos = O.rigid_update(v, st)
    com = np.array([10.0, 10.0, 10.0])
    for k in range(1, 40):
        st, pos = O.rigid_move(v, np.eye(3) * 30.0, vol, [0, 0, 0], (40, 40, 40), st, pos, np.tile(F / len(v), (len(v), 1)))
        vk = F / (m * d) * (1 - (1 - d) ** k)
        com = com + vk
        np.testing.assert_allclose(st[3:6], vk, rtol=1e-12, atol=1e-18)
        np.testing.assert_allclose(st[0:3], com, rtol=0, atol=1e-12)
    assert np.abs(st[6:9]).max() < 1e-15


def test_isotropic_body_spins_at_constant_rate():
    v, t = _sphere()
    axis = np.array([1.0, 2.0, 2.0]) / 3.0
    om = 0.05
    st = O.rigid_state((10.0, 10.0, 10.0), (1, 0, 0, 0), 1.0, 1.0, 0.0, angvel=om * axis)
    pos = O.rigid_update(v, st)
    for k in range(1, 30):
        st, pos = O.rigid_move(v, np.eye(3) * 40.0, 80.0, [0, 0, 0], (40, 40, 40), st, pos, np.zeros_like(pos))
    np.testing.assert_allclose(st[6:9], om * axis, rtol=1e-15)
    ang = 29 * om
    np.testing.assert_allclose(st[9:13], np.r_[np.cos(ang / 2), np.sin(ang / 2) * axis], rtol=0, atol=1e-13)


def test_centre_of_mass_wraps_with_the_vertices():
    v, t = _sphere()
    st = O.rigid_state((19.4, 5.0, 5.0), (1, 0, 0, 0), 1.0, 1.0, 0.0, vel=(0.2, 0.0, 0.0))
    pos = O.rigid_update(v, st)
    st2, pos2 = O.rigid_move(v, np.eye(3) * 40.0, 80.0, [1, 0, 0], (20, 20, 20), st, pos, np.zeros_like(pos))
    assert st2[0] == pytest.approx(19.6 - 20.0, abs=1e-13) and -0.5 <= st2[0] < 19.5
    np.testing.assert_allclose(pos2, pos + (0.2 - 20.0, 0, 0), rtol=0, atol=1e-12)
