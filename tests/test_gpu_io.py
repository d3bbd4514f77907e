"""On-disk formats through the C-ABI against the files the reference's own
classes wrote (tests/golden/io, oracle/ref/ref_io.cpp), and restarts from a
checkpoint.

Bar: particle packs, VTU and gmsh meshes and the lightweight text equal the
reference's bytes; a run restarted from a lattice + particle checkpoint
continues bitwise as the uninterrupted run.
"""
import os
import struct

import numpy as np
import pytest

from conftest import GOLDEN, load_golden

pytestmark = pytest.mark.gpu

MESHES = ["sphere626", "ellipsoid414"]
SHAPE = (40, 36, 44)


def golden_bytes(mesh, name):
    return open(os.path.join(GOLDEN, "io", mesh, name), "rb").read()


def write_checkpoint(folder, packs, name="checkpoint_fsi"):
    """A one-rank checkpoint in the reference's layout
    (src/fsi/ImmersedBoundaryDynamics_io.hh:17-73) holding the given packs."""
    os.makedirs(os.path.join(folder, name), exist_ok=True)
    with open(os.path.join(folder, name + ".plb"), "w") as h:
        h.write("%d\n1\n%s/0.plb\n" % (len(packs), name))
    with open(os.path.join(folder, name, "0.plb"), "wb") as f:
        f.write(struct.pack("<q", len(packs)) + b"".join(packs))


def boundary(kernel="peskin"):
    """Lattice + boundary with both golden shapes in the reference driver's
    library order (shape ids 0 and 1)."""
    from biblbm_amd.lattice import Lattice
    from biblbm_amd.ibm import ImmersedBoundary
    L = Lattice(*SHAPE, periodic=(1, 1, 1), storage="f64", omega=1 / 0.9)
    ib = ImmersedBoundary(L, kernel)
    for m in MESHES:
        sid = ib.add_shape(load_golden("fsi", m, "rest_pos.npy"), load_golden("fsi", m, "tri.npy"))
        ib.set_shape_tag(sid, m)
    return L, ib


@pytest.mark.parametrize("mesh", MESHES)
@pytest.mark.parametrize("kind", ["neohookean", "skalak", "rbc"])
def test_pack_round_trip_and_forces(device, tmp_path, mesh, kind):
    """A reference pack loads, packs back to the same bytes, and the
    membrane forces of the loaded particle equal the forces the reference
    stored in it."""
    pack = golden_bytes(mesh, "pack_%s.bin" % kind)
    write_checkpoint(str(tmp_path), [pack])
    L, ib = boundary()
    ib.load_checkpoint(str(tmp_path))
    ib.init()
    sid, model, pid = ib.describe(0)
    assert sid == MESHES.index(mesh)
    assert model == {"neohookean": 1, "skalak": 0, "rbc": 2}[kind]
    assert pid == {"neohookean": 3, "skalak": 4, "rbc": 5}[kind]
    assert ib.pack(0) == pack
    nv = len(load_golden("fsi", mesh, "pos.npy"))
    ref = np.frombuffer(pack[25:25 + 72 * nv], "<f8").reshape(nv, 9)[:, 6:9]
    ib.compute_forces()
    F = ib.forces()
    assert np.abs(F - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("mesh", MESHES)
def test_vtk_and_gmsh_match_reference(device, tmp_path, mesh):
    write_checkpoint(str(tmp_path), [golden_bytes(mesh, "pack_rbc.bin")])
    L, ib = boundary()
    ib.load_checkpoint(str(tmp_path))
    ib.init()
    ib.write_vtk(0, tmp_path / "p.vtu")
    ib.write_gmsh(0, tmp_path / "p.msh")
    assert (tmp_path / "p.vtu").read_bytes() == golden_bytes(mesh, "rbc.vtu")
    assert (tmp_path / "p.msh").read_bytes() == golden_bytes(mesh, "rbc.msh")


@pytest.mark.parametrize("mesh", MESHES)
def test_lightweight_matches_reference(device, tmp_path, mesh):
    write_checkpoint(str(tmp_path), [golden_bytes(mesh, "pack_rbc.bin"), golden_bytes(mesh, "pack_neohookean.bin")])
    L, ib = boundary()
    ib.load_checkpoint(str(tmp_path))  # kept in id order: 3, then 5
    ib.init()
    assert [ib.describe(p)[2] for p in range(2)] == [3, 5]
    ib.write_lightweight(tmp_path / "objs.txt", 120)
    assert (tmp_path / "objs.txt").read_bytes() == golden_bytes(mesh, "lightweight.txt")


def test_checkpoint_round_trip(device, tmp_path):
    packs = [golden_bytes(m, "pack_%s.bin" % k) for m in MESHES for k in ("rbc", "skalak")]
    # distinct ids: 5, 4 (sphere626) and 15, 14 (ellipsoid414)
    packs = [p[:16] + struct.pack("<q", struct.unpack_from("<q", p, 16)[0] + 10 * (i // 2)) + p[24:]
             for i, p in enumerate(packs)]
    write_checkpoint(str(tmp_path / "a"), packs)
    L, ib = boundary()
    ib.load_checkpoint(str(tmp_path / "a"))
    ib.init()
    ib.save_checkpoint(str(tmp_path / "b"))
    names = open(tmp_path / "b" / "checkpoint_fsi.plb").read().split()
    assert names == ["4", "1", "checkpoint_fsi/0.plb"]
    L2, ib2 = boundary()
    ib2.load_checkpoint(str(tmp_path / "b"))
    ib2.init()
    order = sorted(range(4), key=lambda i: struct.unpack_from("<q", packs[i], 16)[0])
    for p in range(4):
        assert ib2.pack(p) == packs[order[p]]
    with pytest.raises(Exception):
        ib2.load_checkpoint(str(tmp_path / "b"))  # after init


def _pair_setup(storage):
    from biblbm_amd.lattice import Lattice
    from biblbm_amd.ibm import ImmersedBoundary
    from oracle import oracle_py as O
    L = Lattice(*SHAPE, periodic=(1, 0, 1), storage=storage, omega=1 / 0.9)
    rng = np.random.default_rng(3)
    L.put_populations(O.lbm_init_equilibrium(rng.uniform(0.99, 1.01, SHAPE), rng.uniform(-0.02, 0.02, SHAPE + (3,))))
    mask = np.zeros(SHAPE, np.uint8)
    mask[:, 0, :] = mask[:, -1, :] = 1
    L.set_mask(mask)
    L.set_body_force([2e-6, 0, 0])
    ib = ImmersedBoundary(L, "peskin")
    ids = []
    for m in MESHES:
        ids.append(ib.add_shape(load_golden("fsi", m, "rest_pos.npy"), load_golden("fsi", m, "tri.npy")))
    return L, ib, ids


@pytest.mark.parametrize("storage,fused", [("f64", 0), ("f64", 1), ("f16s", 0), ("f32", 1)])
def test_restart_continues_bitwise(device, tmp_path, storage, fused):
    """N steps, checkpoint (lattice + particles), M more steps; a fresh
    lattice and boundary loaded from the checkpoint run the same M steps
    to the same bits."""
    L1, ib1, sids = _pair_setup(storage)
    a_pos = load_golden("fsi", "sphere626", "pos.npy") - [0, 0, 12.0]
    b_pos = load_golden("fsi", "ellipsoid414", "pos.npy") + [0, 0, 10.0]
    ib1.add_particle(sids[0], 2, load_golden("fsi", "sphere626", "rbc_params.npy"), a_pos)
    ib1.add_particle(sids[1], 0, [0.02, 10.0], b_pos)
    ib1.init()
    ib1.set_fused_staging(fused)
    ib1.coupled_steps(7, use_graph=True)
    L1.save_checkpoint(tmp_path / "lbm.bin")
    ib1.save_checkpoint(str(tmp_path))
    ib1.coupled_steps(9, use_graph=True)

    L2, ib2, _ = _pair_setup(storage)
    L2.load_checkpoint(tmp_path / "lbm.bin")
    assert L2.steps == 7
    ib2.load_checkpoint(str(tmp_path))
    ib2.init()
    ib2.set_fused_staging(fused)
    ib2.coupled_steps(9, use_graph=True)
    assert np.array_equal(ib1.positions(), ib2.positions())
    assert np.array_equal(L1.get_populations(), L2.get_populations())
    assert np.array_equal(L1.force(), L2.force())


def test_lattice_checkpoint_rejects_other_sizes(device, tmp_path):
    from biblbm_amd import capi
    from biblbm_amd.lattice import Lattice
    L = Lattice(8, 8, 8, storage="f32")
    L.save_checkpoint(tmp_path / "l.bin")
    with pytest.raises(capi.BlError):
        Lattice(8, 8, 9, storage="f32").load_checkpoint(tmp_path / "l.bin")
    with pytest.raises(capi.BlError):
        Lattice(8, 8, 8, storage="f64").load_checkpoint(tmp_path / "l.bin")
