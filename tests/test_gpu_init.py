"""Initial-condition pipeline through the C-ABI against the reference's
own output (tests/golden/init): the RBC volume calibration on the device
(bl_rbc_shrink_volume: shrink_rbc_volume / relax_nodes,
progs/common/calibrate_rbc_shape.h:13-72, src/fsi/RBCParticle.hh:318-342)
and the host position initializer (bl_positions_*,
src/fsi/ParticlePositionInitializer.hh:18-86).

Bar: calibrated positions within 1e-10 and velocities within 1e-9 of the
reference's (relative to the largest value; the membrane forces agree to
1e-12 per step); positions of the initializer bit-exact."""
import ctypes

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def boundary_with(n_copies, params, shift=30.0):
    from biblbm_amd.lattice import Lattice
    from biblbm_amd.ibm import ImmersedBoundary
    rest = load_golden("fsi", "sphere626", "rest_pos.npy")
    tri = load_golden("fsi", "sphere626", "tri.npy")
    L = Lattice(8, 8, 8, periodic=(0, 0, 0))  # no centre-of-mass remap
    ib = ImmersedBoundary(L, "peskin")
    sid = ib.add_shape(rest, tri)
    for k in range(n_copies):
        ib.add_particle(sid, 2, params, rest + [shift * k, 0, 0])
    ib.init()
    return L, ib, rest


@pytest.mark.parametrize("n_copies", [1, 3])
def test_shrink_rbc_volume_matches_reference(device, n_copies):
    prm = load_golden("init", "shrink626", "rbc_params.npy")
    args = load_golden("init", "shrink626", "args.npy")
    L, ib, rest = boundary_with(n_copies, prm)
    d = ib.shrink_rbc_volume(args[0], int(args[1]), args[2])
    assert np.all(d == args[2])
    pos = load_golden("init", "shrink626", "pos.npy")
    vel = load_golden("init", "shrink626", "vel.npy")
    for k in range(n_copies):
        got = ib.positions(k) - [30.0 * k, 0, 0]
        err = np.abs(got - pos).max() / np.abs(pos).max()
        verr = np.abs(ib.velocities(k) - vel).max() / np.abs(vel).max()
        print(k, "position err", err, "velocity err", verr)
        assert err <= 1e-10 and verr <= 1e-9
        obs = ib.observables(k)
        av = load_golden("init", "shrink626", "area_volume.npy")
        assert abs(obs["area"] - av[0]) <= 1e-10 * av[0] and abs(obs["volume"] - av[1]) <= 1e-10 * av[1]
    # every copy calibrates the same way (up to the rounding of its place)
    if n_copies > 1:
        assert np.abs(ib.positions(0) - (ib.positions(1) - [30.0, 0, 0])).max() <= 1e-10 * np.abs(pos).max()
    # the calibrated particle keeps vol_desired = vol_final (it packs it)
    import struct
    tail = ib.pack(0)[-64:]
    assert abs(struct.unpack("<8d", tail)[7] - args[0]) <= 1e-12 * args[0]


def test_shrink_rbc_volume_reports_divergence(device):
    """A membrane too stiff for the explicit relaxation diverges for every
    damping: the retries double it until 2 and the call fails, as the
    oracle's calibration does (the reference would retry without end)."""
    from biblbm_amd import capi
    from oracle import oracle_py as O
    rest = load_golden("fsi", "sphere626", "rest_pos.npy")
    tri = load_golden("fsi", "sphere626", "tri.npy")
    prm = O.create_rbc_params(rest, tri, 0.2, 2.0, 0.02)
    with pytest.raises(FloatingPointError):
        p = O.Particle(rest, tri, 2, prm, rest.copy())
        with np.errstate(all="ignore"):
            O.shrink_rbc_volume(p, 0.8 * prm[5], 8, 1e-3)
    L, ib, _ = boundary_with(1, prm)
    with pytest.raises(capi.BlError, match="diverges"):
        ib.shrink_rbc_volume(0.8 * prm[5], 8, 1e-3)


def test_position_initializer_matches_reference(device):
    from biblbm_amd.initial import ParticlePositionInitializer, srand
    from biblbm_amd.walls import Boundary
    srand(5)
    init = ParticlePositionInitializer((0, 63, 0, 31, 0, 31), Boundary.pipe(15.5, 15.5, 14.0, 64.0), 3.2)
    assert init.generate_points(40, 3.2, "R")
    assert init.generate_points(30, 1.7, "P")
    assert np.array_equal(init.get_points("R"), load_golden("init", "placement", "R.npy"))
    assert np.array_equal(init.get_points("P"), load_golden("init", "placement", "P.npy"))
    assert ctypes.CDLL(None).rand() == load_golden("init", "placement", "next_rand.npy")[0]
    # larger than the grid allows, and a box with no room left
    assert not init.generate_points(1, 3.3, "X")
    full = ParticlePositionInitializer((0, 8, 0, 8, 0, 8), None, 3.9)
    assert full.generate_points(1, 3.9, "A")
    assert not full.generate_points(1, 3.9, "A")
    assert len(full.get_points("A")) == 1
