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


def _read_vti_vectors(path):
    """First appended Float32 array of a .vti written by VtkImageOutput3D, as
    an (nx, ny, nz, ncomp) array."""
    import re
    raw = open(path, "rb").read()
    head = raw[:raw.index(b"<AppendedData")].decode()
    ext = [int(s) for s in re.search(r'WholeExtent="([^"]+)"', head).group(1).split()]
    ncomp = int(re.search(r'NumberOfComponents="(\d+)"', head).group(1))
    nx, ny, nz = ext[1] - ext[0] + 1, ext[3] - ext[2] + 1, ext[5] - ext[4] + 1
    start = raw.index(b"_", raw.index(b"<AppendedData")) + 1
    nbytes = int(np.frombuffer(raw[start:start + 8], dtype="<u8")[0])
    assert nbytes == nx * ny * nz * ncomp * 4
    a = np.frombuffer(raw[start + 8:start + 8 + nbytes], dtype="<f4").reshape(nz, ny, nx, ncomp)
    return a.transpose(2, 1, 0, 3)


REF_PROG_DIR = os.path.join(ROOT, "biblbm_amd", "examples", "ref_progs")


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_PROG_DIR, "capsule_in_shear")),
                    reason="the reference's capsule_in_shear prog was not built (needs the reference checkout)")
def test_reference_capsule_in_shear_prog_runs(device, tmp_path):
    """The reference's progs/capsule_in_shear/capsule_in_shear.cpp, compiled
    unchanged against include/, run to its end (20 shear times, 401 velocity
    outputs) on a 61^3 box with a radius-6 capsule: Couette walls at -uMax and
    +uMax, a linear profile away from the capsule, the capsule volume held and
    its Taylor deformation settled to a steady value."""
    from biblbm_amd import meshes
    N, Re, Ca, tau = 60, 1.0, 0.15, 0.8
    v, t = meshes.icosphere(3, 6.0)
    meshes.write_gmsh(str(tmp_path / "sphere_r6_TRImesh_1046n_2088el.msh"), v, t)
    r = subprocess.run([os.path.join(REF_PROG_DIR, "capsule_in_shear"), "-N", str(N), "-Re", str(Re), "-Ca", str(Ca)],
                       cwd=tmp_path, capture_output=True, text=True, timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    import re
    # the prog's own figures: uMax = gamma N / 2 and Nt = ceil(20 / gamma)
    u_max = float(re.search(r"uMax:([0-9.eE+-]+)", r.stdout).group(1))
    Nt = int(re.search(r"Iteration 0 / (\d+)", r.stdout).group(1))
    gamma = 2 * u_max / N
    nu = (tau - 0.5) / 3.0
    a = np.sqrt(Re * nu / gamma)
    assert abs(a - 6.0) < 0.1
    assert abs(Nt - 20 / gamma) <= 1.0 + 1e-3 * Nt
    assert "Iteration %d / %d" % (Nt // 50 * 50, Nt) in r.stdout
    out = tmp_path / "spheroid_N60_Ca0_15_tau0_8"
    d = np.loadtxt(out / "deformation_p0.txt")
    assert d.shape == (Nt // 10 + 1, 11) and np.all(np.isfinite(d))
    np.testing.assert_allclose(d[:, 0], gamma * 10 * np.arange(d.shape[0]), rtol=1e-5)
    vol = d[:, 1]
    assert np.max(np.abs(vol / vol[0] - 1)) < 0.02
    # Taylor deformation (L - B) / (L + B) from the inertia tensor of an
    # ellipsoid: the squared semi-axes go as trace/2 - eigenvalue
    D = []
    for row in d[:, 2:]:
        lam = np.linalg.eigvalsh(row.reshape(3, 3))
        ax = np.sqrt(lam.sum() / 2 - lam)
        D.append((ax.max() - ax.min()) / (ax.max() + ax.min()))
    D = np.array(D)
    assert D[0] < 0.02
    assert 0.1 < D[-1] < 0.7
    tail = D[int(0.75 * len(D)):]
    assert tail.max() - tail.min() < 0.03
    # the last velocity output: the walls move at -uMax and +uMax, and the
    # profile is linear across the gap at a column far from the capsule
    files = sorted(out.glob("vtk*.vti"))
    its = np.array([int(f.stem[3:]) for f in files])
    k = its[1] - its[0]
    assert abs(k - Nt / 400.0) <= 1.0
    np.testing.assert_array_equal(its, k * np.arange(len(its)))
    assert its[-1] <= Nt < its[-1] + k
    u = _read_vti_vectors(str(files[-1]))
    np.testing.assert_allclose(u[:, 0, :, 0], -u_max, rtol=1e-5)
    np.testing.assert_allclose(u[:, N, :, 0], u_max, rtol=1e-5)
    y = np.arange(N + 1)
    np.testing.assert_allclose(u[0, :, 0, 0], u_max * (2 * y / N - 1), atol=0.05 * u_max)
