"""The reference's on-disk formats, as its own classes wrote them
(tests/golden/io, oracle/ref/ref_io.cpp), read back with the layout the
library writes: particle packs (src/fsi/ParticleBase.hh:218-224,
src/fsi/DeformableParticle3D.hh:277-287, src/fsi/DeformableCapsule.hh:217-223,
src/fsi/RBCParticle.hh:95-105), the VTU particle mesh
(src/fsi/ParticleBase.hh:270-398), the gmsh mesh (:400-437) and the
checkpoint file names (src/fsi/ImmersedBoundaryDynamics_io.hh:21-37).
Positions, velocities and parameters in the files are the inputs the
driver set, so a misread field (an offset, an order, a size) fails here.
No GPU."""
import os
import struct

import numpy as np
import pytest

from conftest import GOLDEN, load_golden

MESHES = ["sphere626", "ellipsoid414"]


def read_pack(buf):
    """(header, vertex block (nv, 9), tail bytes) of one pack."""
    type_id, shape_id, pid = struct.unpack_from("<qqq", buf, 0)
    vox = buf[24]
    return (type_id, shape_id, pid, vox), buf[25:]


def golden_io(mesh, name):
    return open(os.path.join(GOLDEN, "io", mesh, name), "rb").read()


@pytest.mark.parametrize("mesh", MESHES)
@pytest.mark.parametrize("kind,pid,shear", [("neohookean", 3, 1), ("skalak", 4, 0)])
def test_capsule_pack_layout(mesh, kind, pid, shear):
    pos = load_golden("fsi", mesh, "pos.npy")
    vel = load_golden("io", mesh, "vel.npy")
    nv = len(pos)
    (t, s, i, vox), rest = read_pack(golden_io(mesh, "pack_%s.bin" % kind))
    ids = load_golden("io", mesh, "type_ids.npy")
    assert (t, i, vox) == (ids[0], pid, 1)
    assert s == MESHES.index(mesh)
    block = np.frombuffer(rest[:72 * nv], "<f8").reshape(nv, 9)
    assert np.array_equal(block[:, 0:3], pos)
    assert np.array_equal(block[:, 3:6], vel)
    assert np.abs(block[:, 6:9]).max() > 0  # the particle's own forces
    tail = rest[72 * nv:]
    assert len(tail) == 4 + 4 * 8
    sm, = struct.unpack_from("<i", tail, 0)
    C, kb, nu_s, G = struct.unpack_from("<4d", tail, 4)
    assert sm == shear
    assert (C, kb, nu_s, G) == (10.0, 0.001, 0.5, 0.02)


@pytest.mark.parametrize("mesh", MESHES)
def test_rbc_pack_layout(mesh):
    pos = load_golden("fsi", mesh, "pos.npy")
    nv = len(pos)
    (t, s, i, vox), rest = read_pack(golden_io(mesh, "pack_rbc.bin"))
    assert (t, i) == (load_golden("io", mesh, "type_ids.npy")[1], 5)
    tail = np.frombuffer(rest[72 * nv:], "<f8")
    assert len(tail) == 8
    # parameters as the driver set them: shear_modulus, L0, k_area_global,
    # k_area_local, k_volume, vol_desired, k_bend, theta0
    G, L0, kag, kal, kv, V0, kb, th = load_golden("io", mesh, "rbc_params.npy")
    assert list(tail) == [L0, G, kag, kal, kb, kv, th, V0]
    # the library's own parameter order is that of the fsi goldens
    assert np.array_equal(load_golden("io", mesh, "rbc_params.npy"),
                          np.array([G, L0, kag, kal, kv, V0, kb, th]))


@pytest.mark.parametrize("mesh", MESHES)
def test_vtu_layout(mesh):
    data = golden_io(mesh, "rbc.vtu")
    pos = load_golden("fsi", mesh, "pos.npy")
    tri = load_golden("fsi", mesh, "tri.npy")
    nv, nt = len(pos), len(tri)
    head, _, app = data.partition(b'<AppendedData encoding="raw">\n_')
    assert b'NumberOfPoints="%d" NumberOfCells="%d"' % (nv, nt) in head
    off = 0
    blocks = []
    for size in (12 * nv, 12 * nv, 12 * nt, 4 * nt, 4 * nt):
        n, = struct.unpack_from("<I", app, off)
        assert n == size
        blocks.append(app[off + 4:off + 4 + n])
        off += 4 + n
        assert b'offset="%d"' % (off - 4 - n) in head
    assert np.array_equal(np.frombuffer(blocks[0], "<f4").reshape(nv, 3), pos.astype(np.float32))
    forces = np.frombuffer(golden_io(mesh, "pack_rbc.bin")[25:25 + 72 * nv], "<f8").reshape(nv, 9)[:, 6:9]
    assert np.array_equal(np.frombuffer(blocks[1], "<f4").reshape(nv, 3), forces.astype(np.float32))
    assert np.array_equal(np.frombuffer(blocks[2], "<i4").reshape(nt, 3), tri)
    assert np.array_equal(np.frombuffer(blocks[3], "<i4"), 3 * np.arange(1, nt + 1))
    assert np.all(np.frombuffer(blocks[4], "<i4") == 5)
    assert app[off:] == b"\n  </AppendedData>\n</VTKFile>\n"


@pytest.mark.parametrize("mesh", MESHES)
def test_gmsh_layout(mesh):
    """The writer numbers nodes and elements from 0 (unlike the gmsh reader,
    src/fsi/ParticleShapeFactory.hh:32-83, which expects 1-based node ids)
    and prints coordinates with the stream's default 6 significant digits."""
    lines = open(os.path.join(GOLDEN, "io", mesh, "rbc.msh")).read().split("\n")
    pos = load_golden("fsi", mesh, "pos.npy")
    tri = load_golden("fsi", mesh, "tri.npy")
    assert lines[:4] == ["$MeshFormat", "2.1 0 8", "$EndMeshFormat", "$Nodes"]
    assert int(lines[4]) == len(pos)
    nodes = np.array([[float(x) for x in ln.split()] for ln in lines[5:5 + len(pos)]])
    assert np.array_equal(nodes[:, 0], np.arange(len(pos)))
    assert np.abs(nodes[:, 1:] - pos).max() <= 5e-6 * np.abs(pos).max()
    k = 5 + len(pos)
    assert lines[k:k + 3] == ["$EndNodes", "$Elements", str(len(tri))]
    el = np.array([[int(x) for x in ln.split()] for ln in lines[k + 3:k + 3 + len(tri)]])
    assert np.array_equal(el[:, :6], np.column_stack([np.arange(len(tri)), np.tile([2, 3, 0, 1, 0], (len(tri), 1))]))
    assert np.array_equal(el[:, 6:], tri)
    assert lines[k + 3 + len(tri):] == ["$EndElements"]


def test_checkpoint_names():
    names = open(os.path.join(GOLDEN, "io", "checkpoint", "names.txt")).read().split()
    # header <base>/<name>.plb; the header lists the rank file relative to
    # the base folder; the rank file is written under the base folder
    assert names == ["out/checkpoint_fsi.plb", "checkpoint_fsi/0.plb", "out/checkpoint_fsi/0.plb"]
