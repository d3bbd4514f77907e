"""The FP64 CPU restatement of the D3Q19 step (oracle/lbm.c) against the
reference's own output and against closed forms.

Golden vectors: tests/golden/lbm/*, written by oracle/ref/ref_lbm.cpp from the
reference sources (Palabos MultiBlockLattice3D + GuoExternalForceVOFDynamics,
BounceBack, interpolation velocity boundary) compiled in serial mode.
"""
import numpy as np
import pytest

from conftest import load_golden
from oracle import oracle_py as O

CASES = ["periodic_vof", "shear_interp", "bb_channel"]


def case_inputs(name):
    g = lambda n: load_golden("lbm", name, n + ".npy")
    par = g("params")
    f0 = g("f0")
    nx, ny, nz = f0.shape[:3]
    walls = int(par[4])
    mask = np.zeros((nx, ny, nz), np.uint8)
    if walls == 1:
        mask[:, 0, :] = 1
        mask[:, -1, :] = 1
    elif walls == 2:
        mask[:, 0, :] = 3
        mask[:, -1, :] = 3
    wall_u = np.array([-par[3], 0, 0, par[3], 0, 0])
    return dict(f0=f0, fN=g("fN"), force=g("force"), phase=g("phase"), mask=mask,
                periodic=g("periodic"), omega=par[0], omega_in=par[1], steps=int(par[2]),
                wall_u=wall_u, uN=g("uN"), rhoN=g("rhoN"), walls=walls)


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference(name):
    c = case_inputs(name)
    f = O.lbm_run(c["f0"], c["force"], c["phase"], c["mask"], c["periodic"],
                  c["omega"], c["omega_in"], c["steps"], 1, c["wall_u"])
    fluid = c["mask"] != 1
    scale = np.abs(c["fN"][fluid]).max()
    assert np.abs(f - c["fN"])[fluid].max() <= 1e-13 * scale
    rho, u = O.lbm_moments(f, c["force"], c["mask"], 1, c["wall_u"])
    # every cell, wall cells included (their reported rho / u are fixed values)
    assert np.abs(rho - c["rhoN"]).max() <= 1e-13
    assert np.abs(u - c["uN"]).max() <= 1e-14


def test_planted_guo_sign_error_is_caught():
    """A Guo source with the wrong sign must fail the golden comparison
    (the check is sensitive to the force term, not only to BGK)."""
    c = case_inputs("periodic_vof")
    f_ok = O.lbm_run(c["f0"], c["force"], c["phase"], c["mask"], c["periodic"],
                     c["omega"], c["omega_in"], c["steps"])
    f_bad = O.lbm_run(c["f0"], -c["force"], c["phase"], c["mask"], c["periodic"],
                      c["omega"], c["omega_in"], c["steps"])
    assert np.abs(f_ok - c["fN"]).max() < 1e-15
    assert np.abs(f_bad - c["fN"]).max() > 1e-7


def test_equilibrium_moments_closed_form():
    rng = np.random.default_rng(1)
    rho = rng.uniform(0.9, 1.1, (4, 3, 5))
    u = rng.uniform(-0.08, 0.08, (4, 3, 5, 3))
    f = O.lbm_init_equilibrium(rho, u)
    c, t = np.zeros((19, 3), np.int32), np.zeros(19)
    import ctypes
    cc = (ctypes.c_int * 3 * 19)()
    tt = (ctypes.c_double * 19)()
    O.lib.or_d3q19  # ensure symbol
    from oracle.oracle_py import lib
    lib.or_d3q19.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.or_d3q19(ctypes.cast(cc, ctypes.c_void_p), ctypes.cast(tt, ctypes.c_void_p))
    c = np.array([[cc[i][d] for d in range(3)] for i in range(19)], float)
    t = np.array(list(tt))
    full = f + t                       # f_i = fBar_i + t_i
    assert np.allclose(full.sum(-1), rho, atol=1e-15)
    j = np.einsum("...i,id->...d", full, c)
    assert np.allclose(j, rho[..., None] * u, atol=1e-15)
    Pi = np.einsum("...i,ia,ib->...ab", full, c, c)
    expect = rho[..., None, None] * (np.eye(3) / 3.0 + u[..., :, None] * u[..., None, :])
    assert np.allclose(Pi, expect, atol=1e-15)
    # opposite pairs and weights
    for i in range(1, 10):
        assert np.array_equal(c[i], -c[i + 9])
    assert abs(t.sum() - 1.0) < 1e-15


def test_guo_source_moments():
    """A forced collision conserves mass and changes momentum by exactly
    F (1 + omega (rho - 1) / 2): the equilibrium is taken at j + rho F / 2
    (u = j/rho + F/2) and the Guo source adds (1 - omega/2) F."""
    rho = 1.01
    f = O.lbm_init_equilibrium(np.array([rho]), np.array([[0.02, -0.01, 0.03]]))[0]
    F = np.array([1e-3, -2e-3, 5e-4])
    om = 1.3
    a = O.collide_cell(f, F, 0.0, om, om)
    b = O.collide_cell(f, np.zeros(3), 0.0, om, om)
    fa, fb = a - f, b - f
    c = np.array([[0, 0, 0], [-1, 0, 0], [0, -1, 0], [0, 0, -1], [-1, -1, 0], [-1, 1, 0], [-1, 0, -1],
                  [-1, 0, 1], [0, -1, -1], [0, -1, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0],
                  [1, -1, 0], [1, 0, 1], [1, 0, -1], [0, 1, 1], [0, 1, -1]], float)
    assert abs(fa.sum()) < 1e-16 and abs(fb.sum()) < 1e-16
    assert np.allclose(fa @ c, F * (1 + om * (rho - 1) / 2), rtol=1e-13, atol=0)
    assert np.allclose(fb @ c, 0.0, atol=1e-17)


D3Q19_C = np.array([[0, 0, 0], [-1, 0, 0], [0, -1, 0], [0, 0, -1], [-1, -1, 0], [-1, 1, 0],
                    [-1, 0, -1], [-1, 0, 1], [0, -1, -1], [0, -1, 1], [1, 0, 0], [0, 1, 0],
                    [0, 0, 1], [1, 1, 0], [1, -1, 0], [1, 0, 1], [1, 0, -1], [0, 1, 1], [0, 1, -1]])


def test_pure_streaming_moves_one_link():
    """Streaming only (no-dynamics cells): a single excess population at a
    corner cell lands exactly one lattice vector away, wrapped periodically."""
    nx, ny, nz = 5, 6, 7
    f0 = O.lbm_init_equilibrium(np.ones((nx, ny, nz)), np.zeros((nx, ny, nz, 3)))
    mask = np.full((nx, ny, nz), 2, np.uint8)  # no dynamics: streaming only
    for i in range(1, 19):
        f = f0.copy()
        f[0, 5, 6, i] += 1e-3
        out = O.lbm_run(f, np.zeros((nx, ny, nz, 3)), np.zeros((nx, ny, nz)), mask,
                        [1, 1, 1], 1.0, 1.0, 1)
        idx = np.argwhere(np.abs(out - f0) > 1e-12)
        assert len(idx) == 1
        expect = (np.array([0, 5, 6]) + D3Q19_C[i]) % [nx, ny, nz]
        assert tuple(idx[0][:3]) == tuple(expect) and idx[0][3] == i


def test_mass_momentum_conservation_periodic():
    c = case_inputs("periodic_vof")
    m0 = O.lbm_conserved(c["f0"])
    f = O.lbm_run(c["f0"], c["force"], c["phase"], c["mask"], c["periodic"],
                  c["omega"], c["omega_in"], 3)
    m3 = O.lbm_conserved(f)
    assert abs(m3[0] - m0[0]) < 1e-11
    # momentum grows by about the total force per step (exactly by
    # F (1 + omega (rho-1)/2) per cell; rho stays within 2 % of 1 here)
    dj = m3[1:] - m0[1:]
    Fsum = 3 * c["force"].reshape(-1, 3).sum(0)
    assert np.allclose(dj, Fsum, rtol=0.05, atol=1e-4 * np.abs(c["force"]).sum())
