"""HIP collide-and-stream (libbiblbm.so, through the C-ABI) against the
reference's golden vectors and the FP64 oracle.

Bar: FP64 storage within 1e-12 of the reference populations (relative to
their largest magnitude); FP32 / FP16-shifted storage within the stated
tolerances of the FP64 oracle.
"""
import numpy as np
import pytest

from conftest import load_golden
from oracle import oracle_py as O
from test_oracle_lbm import case_inputs, CASES

pytestmark = pytest.mark.gpu


def make_lattice(c, storage="f64", block=256, xcd=True):
    from biblbm_amd.lattice import Lattice
    nx, ny, nz = c["f0"].shape[:3]
    L = Lattice(nx, ny, nz, periodic=[bool(p) for p in c["periodic"]], storage=storage,
                omega=c["omega"], omega_in=c["omega_in"])
    L.set_launch(block, xcd)
    if c["walls"] == 2:
        L.set_velocity_walls(1, c["wall_u"][:3], c["wall_u"][3:])
    if c["walls"]:
        L.set_mask(c["mask"])
    L.set_force_field(c["force"])
    L.set_phase(c["phase"].astype(np.uint8))
    L.put_populations(c["f0"])
    return L


@pytest.mark.parametrize("name", CASES)
def test_f64_matches_reference(device, name):
    c = case_inputs(name)
    L = make_lattice(c)
    L.collide_and_stream(c["steps"])
    f = L.get_populations()
    fluid = c["mask"] != 1
    scale = np.abs(c["fN"][fluid]).max()
    err = np.abs(f - c["fN"])[fluid].max() / scale
    print(name, "max rel err vs reference", err)
    assert err <= 1e-12
    rho, u = L.moments()
    assert np.abs(rho - c["rhoN"]).max() <= 1e-12
    assert np.abs(u - c["uN"]).max() <= 1e-13


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("steps", [1, 2, 3])
def test_aa_parity_every_step(device, name, steps):
    """After odd and even step counts (both AA layouts) the populations read
    back through the C-ABI equal the oracle's."""
    c = case_inputs(name)
    L = make_lattice(c)
    L.collide_and_stream(steps)
    ref = O.lbm_run(c["f0"], c["force"], c["phase"], c["mask"], c["periodic"], c["omega"],
                    c["omega_in"], steps, 1, c["wall_u"])
    assert np.abs(L.get_populations() - ref).max() <= 1e-15 * 1e3


@pytest.mark.parametrize("block,xcd", [(64, False), (128, True), (192, True), (256, False)])
def test_launch_shape_invariance(device, block, xcd):
    c = case_inputs("periodic_vof")
    L0 = make_lattice(c)
    L1 = make_lattice(c, block=block, xcd=xcd)
    L0.collide_and_stream(5)
    L1.collide_and_stream(5)
    assert np.array_equal(L0.get_populations(), L1.get_populations())


@pytest.mark.parametrize("storage,tol", [("f32", 2e-6), ("f16s", 2e-3)])
@pytest.mark.parametrize("name", CASES)
def test_reduced_storage_error(device, storage, tol, name):
    """FP32 and FP16-shifted storage (FP32 collision): velocity error against
    the FP64 reference after the case's steps, relative to max |u|."""
    c = case_inputs(name)
    L = make_lattice(c, storage)
    L.collide_and_stream(c["steps"])
    rho, u = L.moments()
    fluid = c["mask"] != 1
    err = np.abs(u - c["uN"])[fluid].max() / np.abs(c["uN"][fluid]).max()
    print(storage, name, "max rel velocity error", err)
    assert err <= tol


def test_launch_shape_rejects_oversized_blocks(device):
    from biblbm_amd.capi import BlError
    c = case_inputs("periodic_vof")
    L = make_lattice(c)
    for bad in (32, 320, 512, 1024):
        with pytest.raises(BlError):
            L.set_launch(bad, True)


def test_put_get_roundtrip_both_layouts(device):
    c = case_inputs("bb_channel")
    L = make_lattice(c)
    for k in range(3):
        f = L.get_populations()
        L.put_populations(f)
        assert np.array_equal(L.get_populations(), f)
        L.collide_and_stream(1)


def test_uniform_body_force_path(device):
    """The uniform-force kernel (no field) equals the field kernel holding the
    same value everywhere, and the oracle."""
    from biblbm_amd.lattice import Lattice
    nx, ny, nz = 9, 12, 7
    rng = np.random.default_rng(5)
    rho = rng.uniform(0.99, 1.01, (nx, ny, nz))
    u = rng.uniform(-0.02, 0.02, (nx, ny, nz, 3))
    f0 = O.lbm_init_equilibrium(rho, u)
    g = np.array([2e-5, -1e-5, 3e-6])
    mask = np.zeros((nx, ny, nz), np.uint8)
    mask[:, 0, :] = mask[:, -1, :] = 1
    A = Lattice(nx, ny, nz, periodic=[1, 0, 1], omega=1 / 0.9)
    A.set_mask(mask)
    A.set_body_force(g)
    A.put_populations(f0)
    A.collide_and_stream(7)
    F = np.broadcast_to(g, (nx, ny, nz, 3))
    ref = O.lbm_run(f0, F, np.zeros((nx, ny, nz)), mask, [1, 0, 1], 1 / 0.9, 1 / 0.9, 7)
    fluid = mask == 0
    assert np.abs(A.get_populations() - ref)[fluid].max() < 1e-15
    assert np.allclose(A.force(), F)


def test_conserved_sums(device):
    c = case_inputs("periodic_vof")
    L = make_lattice(c)
    L.collide_and_stream(4)
    ref = O.lbm_run(c["f0"], c["force"], c["phase"], c["mask"], c["periodic"], c["omega"],
                    c["omega_in"], 4)
    got = L.conserved()
    want = O.lbm_conserved(ref)
    assert abs(got[0] - want[0]) <= 1e-12 * want[0]
    assert np.allclose(got[1:], want[1:], rtol=1e-10, atol=1e-14)


def test_taylor_green_decay_large(device):
    """Size-independent property at a full-size lattice (128^3, FP64): mass
    is conserved to round-off and the Taylor-Green velocity decays at
    nu (kx^2 + ky^2) with nu = (tau - 1/2)/3, within 1 %."""
    from biblbm_amd.lattice import Lattice
    n = 128
    tau = 0.8
    u0 = 0.03
    k = 2 * np.pi / n
    x = np.arange(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    u = np.zeros((n, n, n, 3))
    u[..., 0] = u0 * np.cos(k * X) * np.sin(k * Y)
    u[..., 1] = -u0 * np.sin(k * X) * np.cos(k * Y)
    rho = 1.0 - 3 * u0 ** 2 / 4 * (np.cos(2 * k * X) + np.cos(2 * k * Y))
    L = Lattice(n, n, n, storage="f64", omega=1 / tau)
    L.init_equilibrium(rho, u)
    m0 = L.conserved()[0]
    steps = 200
    L.collide_and_stream(steps)
    m1 = L.conserved()[0]
    _, uN = L.moments()
    assert abs(m1 - m0) <= 1e-13 * m0
    nu = (tau - 0.5) / 3
    ratio = np.abs(uN[..., 0]).max() / u0
    expect = np.exp(-2 * nu * k * k * steps)
    assert abs(ratio / expect - 1) < 0.01
