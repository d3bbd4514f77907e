"""The oracle's curved-wall pieces against the reference and against closed
forms.

- Guo wall (wall-node list and populations after N collide-and-stream steps)
  against golden vectors written by the reference's own GuoRigidWallBoundary
  (oracle/ref/ref_guo.cpp): node list exact, populations within 1e-14 of
  their scale on fluid and wall cells.
- Boundaries: plate crossings and normals in closed form; pipe crossings lie
  on the circle; distances signed as the reference's.
- Near-wall blending: a vertex at blend weight >= 1 interpolates / spreads
  exactly as the plain kernel, at weight <= 0 exactly as the top hat, and in
  between as the linear mix of the two computed separately; spreading keeps
  the total force; boundary vertices spread after every bulk vertex.
- Reflection: plate reflection in closed form; pipe reflection keeps the
  vertex inside and the travelled length.
"""
import numpy as np
import pytest

from conftest import load_golden
from oracle import oracle_py as O

GUO_CASES = ["pipe_static", "pipe_moving", "plates_couette"]


def guo_case(name):
    g = lambda n: load_golden("guo", name, n + ".npy")
    prm = g("params")
    kind = int(prm[2])
    b = O.Boundary.pipe(prm[3], prm[4], prm[5], prm[6]) if kind == 0 else O.Boundary.plates(prm[3], prm[4])
    return dict(f0=g("f0"), f1=g("f1"), fN=g("fN"), mask=g("mask"), force=g("force"), periodic=g("periodic"),
                omega=prm[0], steps=int(prm[1]), boundary=b, cell=g("wall_cell"), dir=g("wall_dir"),
                delta=g("wall_delta"), wall_pos=g("wall_pos"), wall_vel=g("wall_vel"))


@pytest.mark.parametrize("name", GUO_CASES)
def test_guo_wall_nodes_match_reference(name):
    c = guo_case(name)
    w = O.guo_wall_nodes(c["f0"].shape[:3], c["boundary"])
    assert np.array_equal(w["cell"], c["cell"])
    assert np.array_equal(w["dir"], c["dir"])
    assert np.array_equal(w["delta"], c["delta"])
    assert np.array_equal(w["wall_pos"], c["wall_pos"])


@pytest.mark.parametrize("name", GUO_CASES)
def test_guo_wall_steps_match_reference(name):
    c = guo_case(name)
    w = O.guo_wall_nodes(c["f0"].shape[:3], c["boundary"])
    shape = c["f0"].shape[:3]
    wall = np.zeros(shape, bool)
    wall[np.unravel_index(c["cell"], shape)] = True
    check = (c["mask"] == 0) | wall
    scale = np.abs(c["fN"][check]).max()
    for steps, ref in ((1, c["f1"]), (c["steps"], c["fN"])):
        f = O.lbm_run(c["f0"], c["force"], np.zeros(shape), c["mask"], c["periodic"], c["omega"], c["omega"],
                      steps, guo=(w, c["wall_vel"]))
        assert np.abs(f - ref)[check].max() <= 1e-14 * scale


def test_guo_no_dynamics_nodes_keep_only_the_non_equilibrium_part():
    """A wall node with no dynamics gets fNeq_w alone: with a uniform fluid
    at equilibrium around it, its populations vanish (to round-off), while
    a fluid wall node gets the equilibrium at the extrapolated momentum."""
    shape = (4, 16, 16)
    b = O.Boundary.pipe(7.5, 7.5, 6.2)
    w = O.guo_wall_nodes(shape, b)
    rho = np.full(shape, 1.01)
    u = np.zeros(shape + (3,))
    u[..., 0] = 0.01
    f = O.lbm_init_equilibrium(rho, u)
    cells = np.unravel_index(w["cell"], shape)
    out = O.guo_wall_apply(f, [1, 0, 0], w, [0.0, 0.0, 0.0], np.full(shape, 2, np.uint8))
    assert np.abs(out[cells]).max() <= 1e-16
    out = O.guo_wall_apply(f, [1, 0, 0], w, [0.0, 0.0, 0.0], np.zeros(shape, np.uint8))
    assert np.abs(out[cells][:, 0]).min() > 1e-3


def test_guo_static_wall_at_delta_half_reproduces_linear_extrapolation():
    """Plates at y0 = 1.5: the wall nodes sit at y = 1 with delta = 0.5.  For
    a linear shear u_x(y) = a (y - y0) at equilibrium the extrapolated wall
    momentum vanishes at the wall velocity 0: j_w = (delta - 1) j1 +
    (1 - delta)/(1 + delta) (delta - 1) j2 = -0.5 j1 - 1/6 j2, evaluated
    independently here."""
    shape = (3, 10, 3)
    b = O.Boundary.plates(1.5, 7.5)
    w = O.guo_wall_nodes(shape, b)
    lo = w["cell"][np.unravel_index(w["cell"], shape)[1] == 1]
    assert len(lo) == 9 and np.allclose(w["delta"][np.isin(w["cell"], lo)], 0.5)
    y = np.arange(shape[1], dtype=float)
    u = np.zeros(shape + (3,))
    u[..., 0] = 0.002 * (y[None, :, None] - 1.5)
    rho = np.ones(shape)
    f = O.lbm_init_equilibrium(rho, u)
    out = O.guo_wall_apply(f, [1, 0, 1], w, [0.0, 0.0, 0.0])
    j1, j2 = 0.002 * (2 - 1.5), 0.002 * (3 - 1.5)
    jw = -0.5 * j1 - (0.5 / 1.5) * 0.5 * j2
    cell = np.unravel_index(lo[0], shape)
    fw = out[cell]
    jx = fw[[10, 13, 14, 15, 16]].sum() - fw[[1, 4, 5, 6, 7]].sum()
    assert abs(jx - jw) < 1e-15


def test_plates_geometry_closed_forms():
    b = O.Boundary.plates(2.0, 9.0)
    t = np.zeros(1)
    x0 = np.array([0.3, 8.5, 1.0])
    x1 = np.array([0.7, 9.5, 1.0])
    assert O.lib.or_boundary_trace_ray(O.ctypes.byref(b), O._p(x0), O._p(x1), O._p(t)) == 1
    assert abs(t[0] - 0.5) < 1e-15
    n = np.zeros(3)
    O.lib.or_boundary_normal(O.ctypes.byref(b), O._p(x0), O._p(n))
    assert np.array_equal(n, [0, 1, 0])
    assert O.lib.or_boundary_distance(O.ctypes.byref(b), O._p(np.array([0, 3.25, 0.]))) == 1.25


def test_pipe_crossings_lie_on_the_circle():
    rng = np.random.default_rng(3)
    b = O.Boundary.pipe(5.0, 6.0, 4.0)
    t = np.zeros(1)
    hits = 0
    for _ in range(200):
        x0 = np.array([0, 5.0, 6.0]) + rng.uniform(-3, 3, 3)
        x1 = x0 + rng.uniform(-4, 4, 3)
        if O.lib.or_boundary_trace_ray(O.ctypes.byref(b), O._p(x0), O._p(x1), O._p(t)):
            p = x0 + t[0] * (x1 - x0)
            assert abs(np.hypot(p[1] - 5.0, p[2] - 6.0) - 4.0) < 1e-12
            n = np.zeros(3)
            O.lib.or_boundary_normal(O.ctypes.byref(b), O._p(p), O._p(n))
            assert abs(np.linalg.norm(n) - 1) < 1e-15 and n[0] == 0
            hits += 1
    assert hits > 20
    inside = np.array([0, 5.0, 7.0])
    assert abs(O.lib.or_boundary_distance(O.ctypes.byref(b), O._p(inside)) - 3.0) < 1e-15


def _field(shape, seed):
    return np.random.default_rng(seed).uniform(-0.05, 0.05, shape + (3,))


@pytest.mark.parametrize("kernel", [0, 1])
def test_near_wall_interpolation_blend(kernel):
    """Weights from distance - half support: >= 1 plain kernel, <= 0 top hat,
    in between (1 - w) top hat + w kernel, each piece from the plain
    interpolation of that kernel."""
    shape = (8, 20, 8)
    per = [1, 0, 1]
    u = _field(shape, 1)
    b = O.Boundary.plates(2.0, 17.0)
    hs = O.lib.or_dirac_half_support(kernel)
    ys = [2.0 + hs + 1.4, 2.0 + hs + 0.999, 2.0 + hs + 0.37, 2.0 + hs - 0.2, 17.0 - hs - 0.6]
    pos = np.array([[3.3, y, 4.1] for y in ys])
    vel = np.empty_like(pos)
    O.lib.or_interpolate_near(kernel, *shape, O._p(np.array(per, np.int32)), O._p(u), len(pos), O._p(pos),
                              O.ctypes.byref(b), O._p(vel))
    plain = O.interpolate(kernel, u, per, pos)
    top = O.interpolate(2, u, per, pos)
    for k, y in enumerate(ys):
        w = (y - 2.0 if y < 9.5 else 17.0 - y) - hs
        near = w < 1
        if not near or w >= 1:
            expect = plain[k]
        elif w <= 0:
            expect = top[k]
        else:
            expect = (1 - w) * top[k] + w * plain[k]
        assert np.abs(vel[k] - expect).max() <= 1e-17 + 1e-15 * np.abs(expect).max()


def test_near_wall_spreading_blend_and_order():
    shape = (8, 20, 8)
    per = np.array([1, 0, 1], np.int32)
    b = O.Boundary.plates(2.0, 17.0)
    kernel = 1
    pos = np.array([[3.3, 9.1, 4.1], [3.6, 4.6, 4.4], [3.3, 9.4, 4.0], [4.1, 3.9, 2.2]])
    F = np.random.default_rng(5).uniform(-1, 1, (4, 3))
    out = np.zeros(shape + (3,))
    O.lib.or_spread_near(kernel, *shape, O._p(per), O._p(out), len(pos), O._p(pos), O._p(F), O.ctypes.byref(b))
    # force conserved (both kernels sum to one inside the domain)
    assert np.abs(out.sum(axis=(0, 1, 2)) - F.sum(axis=0)).max() < 1e-14
    # the same field from the pieces: bulk vertices 0, 2 with the kernel,
    # then vertex 1 (w = 0.6) blended, then vertex 3 (w < 0) top hat
    ref = O.spread(kernel, shape, per, pos[[0, 2]], F[[0, 2]], base=np.zeros(shape + (3,)))
    w = (4.6 - 2.0) - 2.0
    ref = O.spread(kernel, shape, per, pos[[1]], F[[1]] * w, base=ref)
    ref = O.spread(2, shape, per, pos[[1]], F[[1]] * (1 - w), base=ref)
    ref = O.spread(2, shape, per, pos[[3]], F[[3]], base=ref)
    assert np.abs(out - ref).max() < 1e-15


def test_plate_reflection_closed_form():
    b = O.Boundary.plates(1.0, 9.0)
    pos = np.array([[2.0, 8.7, 3.0], [2.0, 5.0, 3.0]])
    vel = np.array([[0.1, 0.5, -0.2], [0.1, 0.2, 0.0]])
    out = pos.copy()
    O.lib.or_move_reflect(2, O._p(out), O._p(vel), O.ctypes.byref(b))
    # upper plate: y' = 2*9 - (8.7 + 0.5); x, z advance freely
    assert np.allclose(out[0], [2.1, 18.0 - 9.2, 2.8], atol=1e-15)
    assert np.array_equal(out[1], pos[1] + vel[1])


def test_pipe_reflection_keeps_vertex_inside_and_path_length():
    b = O.Boundary.pipe(6.0, 6.0, 5.0)
    rng = np.random.default_rng(9)
    for _ in range(100):
        ang = rng.uniform(0, 2 * np.pi)
        r = rng.uniform(4.5, 4.99)
        p = np.array([[1.0, 6.0 + r * np.cos(ang), 6.0 + r * np.sin(ang)]])
        v = np.array([[0.05, 0.6 * np.cos(ang), 0.6 * np.sin(ang)]])
        out = p.copy()
        O.lib.or_move_reflect(1, O._p(out), O._p(v), O.ctypes.byref(b))
        assert np.hypot(out[0, 1] - 6.0, out[0, 2] - 6.0) < 5.0
        # straight to the wall, then the reflected rest: the distance to the
        # impact point adds up to |v|
        x0, x1 = p[0], p[0] + v[0]
        t = np.zeros(1)
        assert O.lib.or_boundary_trace_ray(O.ctypes.byref(b), O._p(x0), O._p(x1), O._p(t))
        hit = x0 + t[0] * v[0]
        assert abs(np.linalg.norm(hit - x0) + np.linalg.norm(out[0] - hit) - np.linalg.norm(v[0])) < 1e-12
