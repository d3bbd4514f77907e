"""The C++ capsule_in_shear prog (include/biblbm_plb.h: MultiBlockLattice3D,
interp velocity boundaries, ImmersedBoundaryDynamics3D and the
applyProcessingFunctional / collideAndStream loop) against the same workload
driven through the Python objects: vertex positions after the same number of
coupled steps.  The prog places the capsule with set_center_of_mass, the
Python workload adds the centre to the mesh, so the two differ by rounding
only."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_capsule_in_shear_prog_matches_python_path(device, tmp_path):
    from biblbm_amd import meshes, workloads
    n, steps, Ca, radius = 32, 20, 0.1, 5.0
    v, t = meshes.icosphere(2, radius)
    msh = tmp_path / "capsule.msh"
    meshes.write_gmsh(str(msh), v, t)
    exe = os.path.join(ROOT, "biblbm_amd", "examples", "capsule_in_shear")
    out = str(tmp_path / "run")
    r = subprocess.run([exe, str(msh), str(n), str(steps), str(Ca), str(radius), out], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    pos = np.loadtxt(out + "_final.txt")
    W = workloads.capsule_in_shear("f64", n=n, radius=radius, level=2, Ca=Ca)
    W.ibm.coupled_steps(steps, use_graph=False)
    ref = W.ibm.positions()
    moved = ref - (v + [n / 2.0, (n - 1) / 2.0, n / 2.0])
    assert np.linalg.norm(pos - ref) <= 1e-8 * np.linalg.norm(moved)
    deform = np.loadtxt(out + "_deformation.txt")
    assert deform.shape[0] == steps // 10 + 1 and np.all(np.isfinite(deform))
    W.ibm.close()
    W.lattice.close()


def test_rbc_in_pipe_prog_matches_python_path(device, tmp_path):
    """The C++ rbc_in_pipe prog with the Guo curved wall, no dynamics beyond
    the pipe and the pipe as the immersed boundary's wall, against the same
    set-up built from the Python objects (GuoWall, outside_mask,
    set_boundary): vertex positions and the velocity profile after the same
    number of coupled steps."""
    from biblbm_amd import meshes
    from biblbm_amd.capi import call
    from biblbm_amd.ibm import ImmersedBoundary
    from biblbm_amd.lattice import Lattice
    from biblbm_amd.walls import Boundary, GuoWall, outside_mask
    R, length, steps = 8.0, 16.0, 12
    v, t = meshes.biconcave(2, 3.5)
    msh = tmp_path / "rbc.msh"
    meshes.write_gmsh(str(msh), v, t)
    exe = os.path.join(ROOT, "biblbm_amd", "examples", "rbc_in_pipe")
    out = str(tmp_path / "run")
    r = subprocess.run([exe, str(msh), str(R), str(length), str(steps), out, "1", "1"], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    pos = np.loadtxt(out + "_final.txt")
    prof = np.loadtxt(out + "_profile.txt")

    c = R + 2
    shape = (17, 21, 21)
    b = Boundary.pipe(c, c, R, length)
    L = Lattice(*shape, periodic=(1, 0, 0), omega=1.0)
    W = GuoWall(L, b)
    W.insert()
    L.set_mask(outside_mask(b, shape, 1.0))
    nu = 0.5 / 3
    L.set_body_force([4 * nu * 0.02 / (R * R), 0, 0])
    L.init_equilibrium(np.ones(shape), np.zeros(shape + (3,)))
    call("bl_lattice_initialize", L.handle)
    vol = np.sum(np.einsum("ij,ij->i", v[t[:, 0]], np.cross(v[t[:, 1]], v[t[:, 2]]))) / 6
    ib = ImmersedBoundary(L, "peskin")
    sid = ib.add_shape(v, t)
    start = v + [length / 2, c, c]
    ib.add_particle(sid, 2, [0.0015, 1.0, 0.5, 0.005, 0.5, vol, 0.0005, 0.0], start)
    ib.init()
    ib.set_boundary(b)
    ib.coupled_steps(steps, use_graph=False)
    ref = ib.positions()
    assert np.linalg.norm(pos - ref) <= 1e-8 * np.linalg.norm(ref - start)
    _, u = L.moments()
    assert np.abs(prof[:, 1] - u[shape[0] // 2, :, int(c), 0]).max() <= 1e-12
    ib.close()
    W.close()
    L.close()


def test_rbc_in_pipe_prog_with_interactions(device, tmp_path):
    """The rbc_in_pipe prog with its input.xml interactions (Morse wall
    decorator on the pipe, Morse particle-particle force) and a cell close
    enough to the wall for the decorator to act, against the Python path
    with the same decorator and force."""
    from biblbm_amd import meshes
    from biblbm_amd.capi import call
    from biblbm_amd.ibm import ImmersedBoundary
    from biblbm_amd.lattice import Lattice
    from biblbm_amd.potentials import MorsePotential
    from biblbm_amd.walls import Boundary, GuoWall, outside_mask
    R, length, steps = 6.0, 16.0, 10
    v, t = meshes.biconcave(2, 4.2)
    msh = tmp_path / "rbc.msh"
    meshes.write_gmsh(str(msh), v, t)
    exe = os.path.join(ROOT, "biblbm_amd", "examples", "rbc_in_pipe")
    out = str(tmp_path / "run")
    r = subprocess.run([exe, str(msh), str(R), str(length), str(steps), out, "1", "1", "f64", "1"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    pos = np.loadtxt(out + "_final.txt")

    c = R + 2
    shape = (17, 17, 17)
    b = Boundary.pipe(c, c, R, length)
    L = Lattice(*shape, periodic=(1, 0, 0), omega=1.0)
    W = GuoWall(L, b)
    W.insert()
    L.set_mask(outside_mask(b, shape, 1.0))
    nu = 0.5 / 3
    L.set_body_force([4 * nu * 0.02 / (R * R), 0, 0])
    L.init_equilibrium(np.ones(shape), np.zeros(shape + (3,)))
    call("bl_lattice_initialize", L.handle)
    vol = np.sum(np.einsum("ij,ij->i", v[t[:, 0]], np.cross(v[t[:, 1]], v[t[:, 2]]))) / 6
    ib = ImmersedBoundary(L, "peskin")
    sid = ib.add_shape(v, t)
    start = v + [length / 2, c, c]
    ib.add_particle(sid, 2, [0.0015, 1.0, 0.5, 0.005, 0.5, vol, 0.0005, 0.0], start)
    ib.init()
    ib.set_boundary(b)
    ib.add_wall_interaction(b, MorsePotential(2.0, 2.0, 1e-4))
    ib.add_pp_force(MorsePotential(1.0, 1.0, 1e-3))
    ib.coupled_steps(steps, use_graph=False)
    ref = ib.positions()
    assert np.linalg.norm(pos - ref) <= 1e-8 * np.linalg.norm(ref - start)
    # the decorator acts on the rim
    dist = R - np.hypot(start[:, 1] - c, start[:, 2] - c)
    assert (dist < 2.0).sum() > 10
    ib.close()
    W.close()
    L.close()


@pytest.mark.parametrize("storage", ["f64", "f16s"])
def test_rbc_in_pipe_prog_restarts_from_checkpoint(device, tmp_path, storage):
    """The prog's checkpoint flow (write_checkpoint every interval:
    parallelIO::save, fsi.save_checkpoint, Time::write_checkpoint, and the
    lightweight particle data; then a run started from those files): the
    resumed run ends on the uninterrupted run's vertex positions and
    velocity profile, to the bit."""
    from biblbm_amd import meshes
    v, t = meshes.biconcave(2, 3.5)
    msh = tmp_path / "rbc.msh"
    meshes.write_gmsh(str(msh), v, t)
    exe = os.path.join(ROOT, "biblbm_amd", "examples", "rbc_in_pipe")
    args = [exe, str(msh), "8", "16", "13"]
    full = subprocess.run(args + ["full", "1", "1", storage, "0", "5", "0"], cwd=tmp_path, capture_output=True,
                          text=True, timeout=120)
    print(full.stdout, full.stderr)
    assert full.returncode == 0
    assert open(tmp_path / "checkpoint_time").read().strip() == "10"
    assert (tmp_path / "checkpoint_fsi.plb").read_text().split() == ["1", "1", "checkpoint_fsi/0.plb"]
    lw = (tmp_path / "objsP0-10.txt").read_text().split("\n")
    assert lw[0] == "iteration:10" and lw[1].startswith("rbc:0:")
    res = subprocess.run(args + ["resumed", "1", "1", storage, "0", "5", "1"], cwd=tmp_path, capture_output=True,
                         text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0
    for suffix in ("_final.txt", "_profile.txt"):
        assert (tmp_path / ("full" + suffix)).read_text() == (tmp_path / ("resumed" + suffix)).read_text()


def test_initial_condition_prog_matches_reference(device, tmp_path):
    """The C++ initial_condition prog (create_rbc_params, shrink_rbc_volume,
    ParticlePositionInitializer, save_checkpoint) on the 626-vertex sphere
    mesh with the parameters of the reference's own calibration and
    placement runs (tests/golden/init): calibrated vertices within 1e-10,
    sphere centres bit-exact, and a checkpoint holding the placed cells."""
    from biblbm_amd import meshes
    from conftest import load_golden
    rest = load_golden("fsi", "sphere626", "rest_pos.npy")
    tri = load_golden("fsi", "sphere626", "tri.npy")
    msh = tmp_path / "rbc.msh"
    meshes.write_gmsh(str(msh), rest, tri)
    args = load_golden("init", "shrink626", "args.npy")
    exe = os.path.join(ROOT, "biblbm_amd", "examples", "initial_condition")
    r = subprocess.run([exe, str(msh), "0.02", "0.4", "0.002", "0.8", str(int(args[1])), "%.17g" % args[2], "5", "40",
                        "ic"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    shape = np.loadtxt(tmp_path / "ic_shape.txt")
    ref = load_golden("init", "shrink626", "pos.npy")
    assert np.abs(shape - ref).max() <= 1e-10 * np.abs(ref).max()
    assert np.array_equal(np.loadtxt(tmp_path / "ic_R.txt"), load_golden("init", "placement", "R.npy"))
    assert np.array_equal(np.loadtxt(tmp_path / "ic_P.txt"), load_golden("init", "placement", "P.npy"))
    assert (tmp_path / "checkpoint_fsi.plb").read_text().split() == ["40", "1", "checkpoint_fsi/0.plb"]
