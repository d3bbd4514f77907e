"""Ensembles of shear boxes (bl_ensemble_*), through the C-ABI."""
import numpy as np
import pytest

from oracle import oracle_py as O

pytestmark = pytest.mark.gpu


def _boxes(sides=(24, 32, 28), storage="f64"):
    from biblbm_amd import workloads
    return [workloads.shear_box(storage, s, tau, Ca, G=0.004, seed=k, radius=0.2 * s)
            for k, (s, tau, Ca) in enumerate(zip(sides, (0.8, 1.1, 0.95), (0.3, 0.5, 0.4)))]


def _close(ws):
    for w in ws:
        w.ibm.close()
        w.lattice.close()


@pytest.mark.parametrize("steps", [6, 5])
def test_ensemble_equals_boxes_stepped_alone(device, steps):
    """One graph for all boxes (each a parallel branch) gives bitwise the
    state of every box stepped on its own, for even and odd step counts."""
    from biblbm_amd.ensemble import Ensemble
    A, B = _boxes(), _boxes()
    ens = Ensemble([(w.lattice, w.ibm) for w in A])
    ens.steps(steps, use_graph=True)
    for w in B:
        w.ibm.coupled_steps(steps, use_graph=False)
    for a, b in zip(A, B):
        assert a.lattice.steps == b.lattice.steps == steps
        assert np.array_equal(a.ibm.positions(), b.ibm.positions())
        assert np.array_equal(a.lattice.get_populations(), b.lattice.get_populations())
    # the captured graph is replayed again after eager work on one box
    A[1].ibm.coupled_steps(1, use_graph=False)
    B[1].ibm.coupled_steps(1, use_graph=False)
    ens.steps(4)
    for w in B:
        w.ibm.coupled_steps(4, use_graph=False)
    for a, b in zip(A, B):
        assert np.array_equal(a.ibm.positions(), b.ibm.positions())
    ens.close()
    _close(A + B)


def test_ensemble_box_matches_oracle(device):
    """A shear box inside an ensemble against the oracle's composed steps
    (regularized velocity walls, biconcave cell with model 3 forces)."""
    from biblbm_amd.ensemble import Ensemble
    ws = _boxes(sides=(20, 24))
    w = ws[0]
    f = w.lattice.get_populations()
    shape = w.lattice.shape
    mask = np.zeros(shape, np.uint8)
    mask[:, 0, :] = mask[:, -1, :] = 3
    u_w = w.params["wall_u"]
    wall_u = np.array([-u_w, 0, 0, u_w, 0, 0])
    from biblbm_amd import meshes, workloads
    R = w.params["radius"]
    v, t = meshes.biconcave(3, R)
    pos0 = w.ibm.positions()
    p = O.Particle(v, t, 3, workloads.skalak_rbc_params(w.params["G"], 10.0, R, meshes.signed_volume(v, t),
                                                        0.02 * w.params["G"] * R * R), pos0)
    ens = Ensemble([(x.lattice, x.ibm) for x in ws])
    steps = 4
    ens.steps(steps)
    force = np.zeros(shape + (3,))
    for _ in range(steps):
        f, force, _ = O.coupled_step(f, force, np.zeros(shape), mask, (1, 0, 1), 1 / w.params["tau"],
                                     1 / w.params["tau"], (0, 0, 0), [p], "peskin", 1, wall_u)
    moved = p.pos - pos0
    err = np.linalg.norm(w.ibm.positions() - p.pos) / np.linalg.norm(moved)
    assert err <= 1e-8
    _, u = w.lattice.moments()
    _, uref = O.lbm_moments(f, force, mask, 1, wall_u)
    assert np.linalg.norm(u - uref) <= 1e-10 * np.linalg.norm(uref)
    ens.close()
    _close(ws)


def test_ensemble_with_lattice_only_box(device):
    from biblbm_amd.ensemble import Ensemble
    from biblbm_amd.lattice import Lattice
    ws = _boxes(sides=(24,))
    L = Lattice(16, 16, 16, periodic=(1, 1, 1), storage="f64", omega=1 / 0.8)
    rng = np.random.default_rng(1)
    f0 = O.lbm_init_equilibrium(rng.uniform(0.99, 1.01, (16, 16, 16)), rng.uniform(-0.02, 0.02, (16, 16, 16, 3)))
    L.put_populations(f0)
    ens = Ensemble([(ws[0].lattice, ws[0].ibm), (L, None)])
    ens.steps(4)
    ref = O.lbm_run(f0, np.zeros((16, 16, 16, 3)), np.zeros((16, 16, 16)), np.zeros((16, 16, 16), np.uint8),
                    [1, 1, 1], 1 / 0.8, 1 / 0.8, 4)
    assert np.abs(L.get_populations() - ref).max() <= 1e-14
    ens.close()
    L.close()
    _close(ws)


def test_ensemble_rejects_mismatched_pairs(device):
    from biblbm_amd import capi
    from biblbm_amd.ensemble import Ensemble
    ws = _boxes(sides=(20, 24))
    with pytest.raises(capi.BlError):
        Ensemble([(ws[0].lattice, ws[1].ibm)])
    with pytest.raises(capi.BlError):
        Ensemble([(ws[0].lattice, ws[0].ibm), (ws[0].lattice, None)])
    _close(ws)


@pytest.mark.parametrize("storage,steps", [("f64", 6), ("f64", 5), ("f32", 6), ("f16s", 4)])
def test_batched_ensemble_equals_branches_and_boxes_alone(device, storage, steps):
    """Every stage of a step as one launch over all boxes (batched) gives
    bitwise the state of the per-box graph branches and of each box stepped
    on its own, over several graph replays and an odd step count (the last
    step eager)."""
    from biblbm_amd.ensemble import Ensemble
    A, B, C = _boxes(storage=storage), _boxes(storage=storage), _boxes(storage=storage)
    ea = Ensemble([(w.lattice, w.ibm) for w in A])
    ea.set_batched(1)
    eb = Ensemble([(w.lattice, w.ibm) for w in B])
    eb.set_batched(0)
    for n in (steps, 4):
        ea.steps(n, use_graph=True)
        eb.steps(n, use_graph=True)
        for w in C:
            w.ibm.coupled_steps(n, use_graph=False)
        assert ea.batched and not eb.batched
        for a, b, c in zip(A, B, C):
            assert a.lattice.steps == b.lattice.steps == c.lattice.steps
            assert np.array_equal(a.ibm.positions(), c.ibm.positions())
            assert np.array_equal(b.ibm.positions(), c.ibm.positions())
            assert np.array_equal(a.lattice.get_populations(), c.lattice.get_populations())
            assert np.array_equal(b.lattice.get_populations(), c.lattice.get_populations())
    ea.close()
    eb.close()
    _close(A + B + C)


def test_batched_mode_refuses_boxes_that_differ(device):
    """Batched-only mode reports boxes it cannot batch (here a box without
    immersed boundary, whose collision has no force field); the default
    mode falls back to graph branches and steps them correctly."""
    from biblbm_amd import capi
    from biblbm_amd.ensemble import Ensemble
    from biblbm_amd.lattice import Lattice
    ws = _boxes(sides=(24,))
    L = Lattice(16, 16, 16, periodic=(1, 1, 1), storage="f64", omega=1 / 0.8)
    L.init_equilibrium()
    ens = Ensemble([(ws[0].lattice, ws[0].ibm), (L, None)])
    ens.set_batched(1)
    with pytest.raises(capi.BlError):
        ens.steps(2)
    ens.set_batched(-1)
    ens.steps(2)
    assert not ens.batched
    with pytest.raises(capi.BlError):
        ens.set_batched(2)
    ens.close()
    L.close()
    _close(ws)


def test_ensemble_tuner_keeps_the_trajectory(device):
    """The ensemble tuner's trials (batched collide workgroup orders and
    nontemporal levels) are real steps that give the values of the boxes
    stepped alone."""
    from biblbm_amd.ensemble import Ensemble
    A, B = _boxes(), _boxes()
    ens = Ensemble([(w.lattice, w.ibm) for w in A])
    rep = ens.tune(2)
    assert rep["batched"] == 1
    assert rep["best"]["even"] in (0, 1, 4, 16, 64, 256) and rep["best"]["odd"] in (0, 1, 4, 16, 64, 256)
    assert rep["best"]["nt_even"] in (0, 1, 2) and rep["best"]["nt_odd"] in (0, 1, 2)
    assert len(rep["trials"]) == 2 * 6 + 2 * 2  # orders, then the two other levels, per parity
    n = rep["steps_run"]
    assert A[0].lattice.steps == n
    ens.steps(4)
    for w in B:
        w.ibm.coupled_steps(n + 4, use_graph=False)
    for a, b in zip(A, B):
        assert np.array_equal(a.ibm.positions(), b.ibm.positions())
        assert np.array_equal(a.lattice.get_populations(), b.lattice.get_populations())
    ens.close()
    _close(A + B)


@pytest.mark.parametrize("even,odd", [((256, 2), (0, 1)), ((16, 1), (64, 2))])
def test_ensemble_set_launch_keeps_the_trajectory(device, even, odd):
    """Fixed batched collide orders and nontemporal levels
    (bl_ensemble_set_launch), changed between graph runs, give bitwise the
    boxes stepped alone; out-of-range settings are refused."""
    from biblbm_amd import capi
    from biblbm_amd.ensemble import Ensemble
    A, B = _boxes(), _boxes()
    ens = Ensemble([(w.lattice, w.ibm) for w in A])
    ens.set_launch(0, *even)
    ens.set_launch(1, *odd)
    ens.steps(4)
    assert ens.batched
    ens.set_launch(0, 0, 0)
    ens.steps(2)
    for bad in ((2, 0, 0), (0, 3, 0), (0, 6, 0), (1, 0, 3), (0, -1, 0)):
        with pytest.raises(capi.BlError):
            ens.set_launch(*bad)
    for w in B:
        w.ibm.coupled_steps(6, use_graph=False)
    for a, b in zip(A, B):
        assert np.array_equal(a.ibm.positions(), b.ibm.positions())
        assert np.array_equal(a.lattice.get_populations(), b.lattice.get_populations())
    ens.close()
    _close(A + B)
