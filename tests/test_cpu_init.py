"""The initial-condition pipeline's oracle against the reference's own
output (tests/golden/init, oracle/ref/ref_init.cpp): create_rbc_params
(progs/common/calibrate_rbc_shape.h:75-112), shrink_rbc_volume with
relax_nodes (progs/common/calibrate_rbc_shape.h:13-72,
src/fsi/RBCParticle.hh:318-342) and ParticlePositionInitializer
(src/fsi/ParticlePositionInitializer.hh:18-86).  No GPU."""
import ctypes

import numpy as np
import pytest

from conftest import load_golden
from oracle import oracle_py as O


def sphere():
    return load_golden("fsi", "sphere626", "rest_pos.npy"), load_golden("fsi", "sphere626", "tri.npy")


def test_create_rbc_params_matches_reference():
    rest, tri = sphere()
    ref = load_golden("init", "shrink626", "rbc_params.npy")
    args = load_golden("init", "shrink626", "args.npy")
    got = O.create_rbc_params(rest, tri, args[3], args[4], 0.002)
    assert np.allclose(got, ref, rtol=1e-14, atol=0)
    # theta0 of the 626-vertex mesh: the reference run's log
    # (progs/rbc_in_pipe/log_result2.txt:27)
    assert abs(ref[7] - 0.108354) < 5e-7
    from biblbm_amd.initial import create_rbc_params
    assert np.allclose(create_rbc_params(rest, tri, args[3], args[4], 0.002), ref, rtol=1e-14, atol=0)


def shrink(params=None, vol_final=None, damping=None):
    rest, tri = sphere()
    args = load_golden("init", "shrink626", "args.npy")
    prm = load_golden("init", "shrink626", "rbc_params.npy") if params is None else params
    p = O.Particle(rest, tri, 2, prm, rest.copy())
    d = O.shrink_rbc_volume(p, args[0] if vol_final is None else vol_final, int(args[1]),
                            args[2] if damping is None else damping)
    return p, d


def test_shrink_rbc_volume_matches_reference():
    p, d = shrink()
    pos = load_golden("init", "shrink626", "pos.npy")
    vel = load_golden("init", "shrink626", "vel.npy")
    rest = sphere()[0]
    assert d == load_golden("init", "shrink626", "args.npy")[2]
    assert np.abs(pos - rest).max() > 1.0  # the cell did deflate
    assert np.abs(p.pos - pos).max() <= 1e-11 * np.abs(pos).max()
    assert np.abs(p.vel - vel).max() <= 1e-10 * np.abs(vel).max()
    av = load_golden("init", "shrink626", "area_volume.npy")
    assert abs(p.area - av[0]) <= 1e-12 * av[0] and abs(p.volume - av[1]) <= 1e-12 * av[1]


@pytest.mark.parametrize("mistake", ["no_damping", "no_ramp"])
def test_planted_calibration_mistakes_are_caught(mistake):
    """The same comparison fails for an oracle without the damping term or
    with the volume set to its final value from the first step."""
    pos = load_golden("init", "shrink626", "pos.npy")
    args = load_golden("init", "shrink626", "args.npy")
    if mistake == "no_damping":
        p, _ = shrink(damping=1e-30)
    else:
        prm = load_golden("init", "shrink626", "rbc_params.npy").copy()
        rest, tri = sphere()
        p = O.Particle(rest, tri, 2, prm, rest.copy())
        # a ramp of one step: vol_desired is vol_final from the start
        O.shrink_rbc_volume(p, args[0], 1, args[2])
    assert np.abs(p.pos - pos).max() > 1e-6 * np.abs(pos).max()


def test_position_initializer_matches_reference():
    pts = O.place_spheres((0, 63, 0, 31, 0, 31), O.Boundary.pipe(15.5, 15.5, 14.0, 64.0), 3.2,
                          [(40, 3.2), (30, 1.7)], seed=5)
    assert np.array_equal(pts[0], load_golden("init", "placement", "R.npy"))
    assert np.array_equal(pts[1], load_golden("init", "placement", "P.npy"))
    # the same number of draws as the reference
    assert ctypes.CDLL(None).rand() == load_golden("init", "placement", "next_rand.npy")[0]
    # every sphere inside the pipe and apart from the others
    allp = np.concatenate(pts)
    r = np.array([3.2] * 40 + [1.7] * 30)
    d = np.sqrt(((allp[:, None] - allp[None]) ** 2).sum(-1))
    np.fill_diagonal(d, np.inf)
    assert np.all(d >= r[:, None] + r[None])
    assert np.all(np.hypot(allp[:, 1] - 15.5, allp[:, 2] - 15.5) < 14.0 - r)
