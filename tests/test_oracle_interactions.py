"""The oracle's interaction forces against the reference and against closed
forms.

- Collision handler: forces on every node equal, bit for bit, the ones the
  reference's own CollisionHandler left (oracle/ref/ref_pp.cpp: 8 cases over
  closed, partly and fully periodic boxes, Morse and spring potentials, with
  and without the handler's wall forces, nodes in the envelope and outside
  the collision box).
- Wall interaction decorator: equal, bit for bit, to the reference's
  WallInteraction::apply_force on a capsule near plates and inside a pipe.
- Closed forms: the potentials' forces; two nodes in forward-neighbour
  cells repel along their axis with F(d); pairs in cells related by
  (+1, -1, 0) do not interact (the handler visits 7 forward neighbours, not
  13); the pair forces sum to zero (each pair adds +df and -df).
- A textbook all-pairs sum over the same nodes differs from the reference on
  the goldens, so the fixtures do pin the handler's own pair rules.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_golden
from oracle import oracle_py as O

PP_CASES = sorted(d for d in os.listdir(os.path.join(GOLDEN, "pp")) if not d.startswith("wall_"))
WALL_CASES = sorted(d[5:] for d in os.listdir(os.path.join(GOLDEN, "pp")) if d.startswith("wall_"))


def potential_of(kind, p):
    return O.Potential.morse(p[0], p[1], p[2]) if int(kind) == 0 else O.Potential.spring(p[0], p[1])


def pp_case(name):
    g = lambda n: load_golden("pp", name, n + ".npy")
    P = g("params")
    wk = int(P[10])
    wall = None if wk < 0 else (O.Boundary.pipe(P[11], P[12], P[13], P[14]) if wk == 0 else
                                O.Boundary.plates(P[11], P[12]))
    return dict(pos=g("pos"), obj=g("obj"), force=g("force"), u=potential_of(P[0], P[1:4]),
                shape=tuple(int(x) for x in P[4:7]), periodic=[int(x) for x in P[7:10]], wall=wall)


def wall_case(name):
    g = lambda n: load_golden("pp", "wall_" + name, n + ".npy")
    P = g("params")
    b = O.Boundary.pipe(*P[5:9]) if int(P[4]) == 0 else O.Boundary.plates(P[5], P[6])
    return dict(pos=g("pos"), force=g("force"), u=potential_of(P[0], P[1:4]), boundary=b)


@pytest.mark.parametrize("name", PP_CASES)
def test_collision_forces_match_reference(name):
    c = pp_case(name)
    f, pairs = O.collision_forces(c["shape"], c["periodic"], c["u"], c["pos"], c["obj"], c["wall"])
    assert pairs > 0
    assert np.array_equal(f, c["force"])


@pytest.mark.parametrize("name", WALL_CASES)
def test_wall_interaction_matches_reference(name):
    c = wall_case(name)
    f = O.wall_interaction(c["boundary"], c["u"], c["pos"])
    assert np.count_nonzero(np.abs(c["force"]).sum(1)) > 10
    assert np.array_equal(f, c["force"])


def test_potential_closed_forms():
    m = O.Potential.morse(1.5, 2.0, 1e-3)
    for d in (0.0, 0.3, 1.0, 1.5, 2.5):
        x = np.exp(2.0 * (1.5 - d))
        assert O.potential_force(m, d) == pytest.approx(2 * 2.0 * 1e-3 * x * (x - 1), rel=1e-15)
    assert O.potential_force(m, 1.5) == 0.0                       # zero at r_min
    assert O.potential_force(m, -0.7) == O.potential_force(m, 0.0)  # negative distances clamp to 0
    s = O.Potential.spring(1.0, 0.05)
    assert O.potential_force(s, 0.25) == pytest.approx(0.75 * 0.05, rel=1e-15)
    assert O.potential_force(s, 1.0) == 0.0


def _two(p0, p1, u, shape=(8, 8, 8), periodic=(0, 0, 0)):
    pos = np.array([p0, p1], float)
    f, n = O.collision_forces(shape, list(periodic), u, pos, np.array([0, 1], np.int32))
    return f, n


def test_two_nodes_repel_along_their_axis():
    u = O.Potential.spring(1.0, 0.1)
    # cells differ by (1, 1, 0): a forward neighbour
    a, b = np.array([3.3, 3.4, 4.0]), np.array([3.7, 3.6, 4.1])
    f, n = _two(a, b, u)
    assert n == 1
    d = np.linalg.norm(b - a)
    e = (b - a) / d
    F = O.potential_force(u, d)
    assert np.allclose(f[0], -F * e, rtol=0, atol=1e-17)
    assert np.allclose(f[1], F * e, rtol=0, atol=1e-17)


def test_cells_related_by_plus_minus_do_not_interact():
    """Grid cells are [k-0.5, k+0.5): the two nodes sit in cells (i, j+1)
    and (i+1, j), 0.14 apart; neither is a forward neighbour of the other,
    so the handler leaves them alone, as the reference does."""
    u = O.Potential.spring(1.0, 0.1)
    f, n = _two([3.45, 3.55, 4.0], [3.55, 3.45, 4.0], u)
    assert n == 0 and not f.any()
    # the same pair shifted into one cell interacts
    f, n = _two([3.2, 3.3, 4.0], [3.3, 3.2, 4.0], u)
    assert n == 1 and f.any()


@pytest.mark.parametrize("name", PP_CASES)
def test_pair_forces_sum_to_zero(name):
    c = pp_case(name)
    if c["wall"] is not None:
        pytest.skip("wall forces are external")
    f, _ = O.collision_forces(c["shape"], c["periodic"], c["u"], c["pos"], c["obj"])
    assert np.abs(f.sum(0)).max() <= 1e-15 * max(np.abs(f).max(), 1e-30) * len(f)


def _all_pairs(c):
    """Every pair of different owners within the cut-off, minimal image on
    periodic axes, among nodes inside the collision box."""
    pos, obj, u = c["pos"], c["obj"], c["u"]
    cut = u.p[0]
    ext = np.array(c["shape"], float)
    inside = np.all((pos > -0.5 - cut) & (pos < ext - 0.5 + cut), axis=1)
    f = np.zeros_like(pos)
    for a in range(len(pos)):
        for b in range(a + 1, len(pos)):
            if obj[a] == obj[b] or not (inside[a] and inside[b]):
                continue
            dr = pos[b] - pos[a]
            for d in range(3):
                if c["periodic"][d]:
                    dr[d] -= np.floor(dr[d] / ext[d] + 0.5) * ext[d]
            dist = np.linalg.norm(dr)
            if dist <= cut:
                df = (-O.potential_force(u, dist) / dist) * dr
                f[a] += df
                f[b] -= df
    return f


def test_textbook_all_pairs_differs_from_the_reference_rules():
    diffs = 0
    for name in PP_CASES:
        c = pp_case(name)
        if c["wall"] is not None:
            continue
        scale = np.abs(c["force"]).max()
        if np.abs(_all_pairs(c) - c["force"]).max() > 1e-6 * scale:
            diffs += 1
    assert diffs >= 3
