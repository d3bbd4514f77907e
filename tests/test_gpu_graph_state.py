"""Captured coupled-step graphs follow every setter that changes what a
captured step bakes in (kernel arguments such as omega, body force and
buffer pointers, launch choices, interactions).  Each case steps one copy
eagerly and one through graph replay and requires bitwise equality after
the state changed between two graph runs."""
import numpy as np
import pytest

from test_gpu_ibm import _build_pair

pytestmark = pytest.mark.gpu


def _same(L1, ib1, L2, ib2):
    assert L1.steps == L2.steps
    assert np.array_equal(ib1.positions(), ib2.positions())
    assert np.array_equal(L1.get_populations(), L2.get_populations())


def test_lattice_setters_between_graph_runs(device):
    L1, ib1, *_ = _build_pair("peskin")
    L2, ib2, *_ = _build_pair("peskin")
    mask = np.zeros(L1.shape, np.uint8)
    mask[:, 0, :] = mask[:, -1, :] = 1

    def both(fn, n):
        fn(L1)
        fn(L2)
        ib1.coupled_steps(n, use_graph=False)
        ib2.coupled_steps(n, use_graph=True)
        _same(L1, ib1, L2, ib2)

    both(lambda L: None, 6)
    both(lambda L: L.set_body_force((3e-6, -1e-6, 0.5e-6)), 6)
    both(lambda L: L.set_dynamics(1 / 0.8, 1 / 0.8), 6)
    both(lambda L: L.set_mask(mask), 4)
    # the mask buffer is freed: a stale graph would read freed memory
    both(lambda L: L.set_mask(None), 4)
    both(lambda L: L.set_launch(128, False), 4)
    # long runs (the 32-step captured run) after a change as well
    both(lambda L: L.set_body_force((1e-6, 0, 0)), 36)


def test_boundary_setters_between_graph_runs(device):
    from biblbm_amd.potentials import MorsePotential
    from biblbm_amd.walls import Boundary
    L1, ib1, *_ = _build_pair("roma")
    L2, ib2, *_ = _build_pair("roma")
    ib1.coupled_steps(4, use_graph=False)
    ib2.coupled_steps(4, use_graph=True)
    plates = Boundary.plates(0.5, L1.shape[1] - 1.5)
    for ib in (ib1, ib2):
        ib.add_wall_interaction(plates, MorsePotential(12.0, 0.5, 2e-4))
    ib1.coupled_steps(4, use_graph=False)
    ib2.coupled_steps(4, use_graph=True)
    _same(L1, ib1, L2, ib2)
    # the decorator acts: the same steps without it end elsewhere
    L3, ib3, *_ = _build_pair("roma")
    ib3.coupled_steps(8, use_graph=False)
    assert not np.array_equal(ib3.positions(), ib2.positions())
    for ib in (ib1, ib2):
        ib.add_pp_force(MorsePotential(1.0, 1.0, 1e-3))
        ib.set_layered(0)
    ib1.coupled_steps(4, use_graph=False)
    ib2.coupled_steps(4, use_graph=True)
    _same(L1, ib1, L2, ib2)


def test_ensemble_follows_changes_after_capture(device):
    """Interactions added and lattice settings changed on a box after an
    ensemble graph run reach the next graph run."""
    from biblbm_amd.ensemble import Ensemble
    from biblbm_amd.potentials import MorsePotential
    from biblbm_amd.walls import Boundary
    from test_gpu_ensemble import _boxes, _close
    A, B = _boxes(), _boxes()
    ens = Ensemble([(w.lattice, w.ibm) for w in A])
    ens.steps(4, use_graph=True)
    for w in B:
        w.ibm.coupled_steps(4, use_graph=False)
    s0 = A[0].lattice.shape[1]
    for W in (A, B):
        W[0].ibm.add_wall_interaction(Boundary.plates(0.5, s0 - 1.5), MorsePotential(0.4 * s0, 0.5, 2e-4))
        W[1].lattice.set_dynamics(1 / 0.75, 1 / 0.75)
        W[2].lattice.set_velocity_walls(1, [-0.01, 0, 0], [0.01, 0, 0])
    ens.steps(4, use_graph=True)
    for w in B:
        w.ibm.coupled_steps(4, use_graph=False)
    for a, b in zip(A, B):
        _same(a.lattice, a.ibm, b.lattice, b.ibm)
    ens.close()
    _close(A + B)
