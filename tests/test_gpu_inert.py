"""Inert-cell skipping (bl_lattice_set_skip_inert), through the C-ABI: the
set of inert cells against a direct count on the mask, and every cell that
is not inert, the moments, the conservation sums and coupled immersed-
boundary steps bitwise as without skipping."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _velocities():
    """D3Q19 in the reference's order (i and i + 9 opposite)."""
    v = [(0, 0, 0), (-1, 0, 0), (0, -1, 0), (0, 0, -1), (-1, -1, 0), (-1, 1, 0), (-1, 0, -1), (-1, 0, 1),
         (0, -1, -1), (0, -1, 1)]
    return np.array(v + [tuple(-np.array(c)) for c in v[1:]])


def inert_cells(mask, periodic):
    """Bounce-back cells whose 18 neighbours are all bounce-back or past a
    non-periodic face, by direct count."""
    nx, ny, nz = mask.shape
    out = np.zeros(mask.shape, bool)
    for x, y, z in zip(*np.nonzero(mask == 1)):
        ok = True
        for c in _velocities()[1:]:
            p = np.array([x, y, z]) + c
            inside = True
            for a, n in enumerate((nx, ny, nz)):
                if p[a] < 0 or p[a] >= n:
                    if periodic[a]:
                        p[a] %= n
                    else:
                        inside = False
            if inside and mask[tuple(p)] != 1:
                ok = False
                break
        out[x, y, z] = ok
    return out


def _pipe(R=7, length=20):
    from biblbm_amd import workloads
    mask, _, _ = workloads.pipe_mask(R, length)
    return mask


@pytest.mark.parametrize("storage,periodic", [("f64", (1, 0, 0)), ("f32", (1, 1, 0)), ("f16s", (0, 0, 0))])
def test_skip_inert_steps_every_other_cell_bitwise(device, storage, periodic):
    from biblbm_amd.lattice import Lattice
    mask = _pipe()
    shape = mask.shape
    expect = inert_cells(mask, periodic)
    rng = np.random.default_rng(2)
    rho = rng.uniform(0.99, 1.01, shape)
    u = rng.uniform(-0.02, 0.02, shape + (3,))
    runs = []
    for skip in (0, 1):
        L = Lattice(*shape, periodic=periodic, storage=storage, omega=1 / 0.9)
        L.set_mask(mask)
        L.init_equilibrium(rho, u)
        L.set_body_force((2e-5, 0, 0))
        n = L.set_skip_inert(skip)
        assert n == (int(expect.sum()) if skip else 0)
        for p in (0, 1):
            L.set_launch_parity(p, 64 if p else 128, 4, p)
        L.collide_and_stream(7)
        runs.append((L.get_populations(), L.moments(), L.conserved()))
        L.close()
    assert expect.sum() > 0.2 * mask.size
    (fa, (ra, ua), ca), (fb, (rb, ub), cb) = runs
    keep = ~expect
    assert np.array_equal(fa[keep], fb[keep])
    assert np.array_equal(ra, rb) and np.array_equal(ua, ub)
    assert np.array_equal(np.asarray(ca), np.asarray(cb))


def test_skip_inert_follows_the_mask(device):
    """A new mask rebuilds the inert set; turning skipping off clears it."""
    from biblbm_amd.lattice import Lattice
    mask = _pipe()
    L = Lattice(*mask.shape, periodic=(1, 0, 0), storage="f64", omega=1.0)
    L.set_mask(mask)
    assert L.set_skip_inert(1) == int(inert_cells(mask, (1, 0, 0)).sum())
    m2 = mask.copy()
    m2[:, :3, :] = 1
    L.set_mask(m2)
    from biblbm_amd import capi
    import ctypes
    n = ctypes.c_longlong()
    capi.call("bl_lattice_inert_cells", L.handle, ctypes.byref(n))
    assert n.value == int(inert_cells(m2, (1, 0, 0)).sum())
    assert L.set_skip_inert(0) == 0
    with pytest.raises(capi.BlError):
        capi.call("bl_lattice_set_skip_inert", L.handle, 2)
    L.close()


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_skip_inert_in_coupled_pipe_steps(device, storage):
    """A red cell in a bounce-back pipe: vertex positions and fluid cells
    bitwise with and without skipping, eagerly and from captured graphs."""
    from biblbm_amd import workloads
    ws = []
    for skip in (0, 1):
        w = workloads.rbc_in_pipe(storage, R=12, length=48, level=3, cell_radius=6.0)
        n = w.lattice.set_skip_inert(skip)
        assert (n == 0) == (skip == 0)
        w.ibm.coupled_steps(3, use_graph=False)
        w.ibm.coupled_steps(6, use_graph=True)
        ws.append(w)
    a, b = ws
    assert np.array_equal(a.ibm.positions(), b.ibm.positions())
    _, ua = a.lattice.moments()
    _, ub = b.lattice.moments()
    assert np.array_equal(ua, ub)
    for w in ws:
        w.ibm.close()
        w.lattice.close()


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_skip_inert_with_a_capsule_at_the_wall(device, storage):
    """A capsule next to the pipe wall whose 4-point stencils reach inert
    cells two layers into the bounce-back wall, with the velocity staging
    carried by the collision (an inert cell still stages its zero
    velocity): positions and flow bitwise with and without skipping."""
    from biblbm_amd import meshes, workloads
    from biblbm_amd.ibm import ImmersedBoundary, CAPSULE_SKALAK
    from biblbm_amd.lattice import Lattice
    R, length = 10, 24
    mask, _, c = workloads.pipe_mask(R, length)
    inert = inert_cells(mask, (1, 0, 0))
    v, t = meshes.icosphere(2, 3.0)
    pos0 = v + [length / 2, c + R - 3.4, c]
    hits = 0
    for p in pos0:
        b = np.floor(p).astype(int) - 1
        for i, j, k in np.ndindex(4, 4, 4):
            q = ((b[0] + i) % length, b[1] + j, b[2] + k)
            hits += bool(inert[q])
    assert hits > 0  # the case under test: stencil nodes on inert cells
    out = []
    for skip in (0, 1):
        L = Lattice(length, mask.shape[1], mask.shape[2], periodic=(1, 0, 0), storage=storage, omega=1 / 0.9)
        L.set_mask(mask)
        L.set_body_force([1e-5, 0, 0])
        L.init_equilibrium()
        ib = ImmersedBoundary(L, "peskin")
        sid = ib.add_shape(v, t)
        ib.add_particle(sid, CAPSULE_SKALAK, [0.01, 10.0], pos0)
        ib.init()
        ib.set_fused_staging(1)
        n = L.set_skip_inert(skip)
        ib.coupled_steps(3, use_graph=False)
        ib.coupled_steps(4, use_graph=True)
        out.append((ib.positions(), L.moments()[1], n))
        ib.close()
        L.close()
    assert out[0][2] == 0 and out[1][2] == int(inert.sum())
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
