"""The autotuner (bl_tune_*), through the C-ABI: the collide tuner's trials
leave the lattice as it was and the tuned per-parity shapes step bitwise
like the default; the coupled-step tuner's layouts give bitwise the same
trajectory; the immersed-boundary launch parameters keep the trajectory
bitwise (the gather lane count to rounding) and their tuner's choice is
persisted; the tuning table survives a save / load and a lattice of the
tuned size starts with the tuned shapes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _box(s=32, seed=0):
    from biblbm_amd import workloads
    return workloads.shear_box("f64", s, 0.9, 0.3, G=0.004, seed=seed, radius=0.2 * s)


def test_collide_tuner_keeps_state_and_results(device):
    from biblbm_amd import tuning
    a, b = _box(), _box()
    a.ibm.coupled_steps(3, use_graph=False)
    b.ibm.coupled_steps(3, use_graph=False)
    before = a.lattice.get_populations()
    rep = tuning.tune_collide(a.lattice)
    assert len(rep["even"]) == 64 and len(rep["odd"]) == 96
    for par in ("even", "odd"):
        assert rep["best"][par]["block"] in (64, 128, 192, 256)
        assert rep["best"][par]["xcd_remap"] in (0, 1, 3, 4, 16, 64, 256, 1024)
    assert rep["best"]["even"]["nontemporal"] in (0, 1) and rep["best"]["odd"]["nontemporal"] in (0, 1, 2)
    # the pick is the fastest of the three sweep leaders timed again
    for par in ("even", "odd"):
        ref = rep["refined"][par]
        assert len(ref) == 3 and rep["best"][par] == min(ref, key=lambda r: r["ms"])
    assert a.lattice.steps == 3
    assert np.array_equal(a.lattice.get_populations(), before)
    got = [a.lattice.launch(p) for p in (0, 1)]
    assert got[0] == {k: rep["best"]["even"][k] for k in got[0]}
    assert got[1] == {k: rep["best"]["odd"][k] for k in got[1]}
    a.ibm.coupled_steps(6, use_graph=True)
    b.ibm.coupled_steps(6, use_graph=True)
    assert np.array_equal(a.lattice.get_populations(), b.lattice.get_populations())
    assert np.array_equal(a.ibm.positions(), b.ibm.positions())
    for w in (a, b):
        w.ibm.close()
        w.lattice.close()


@pytest.mark.parametrize("storage", ["f64", "f32", "f16s"])
def test_nontemporal_hints_are_bitwise(device, storage):
    """Nontemporal population loads and stores on both parities (level 1: the
    odd step's misaligned z +-1 accesses included; level 2: its z-aligned
    planes only) leave the results unchanged."""
    from biblbm_amd.lattice import Lattice
    rng = np.random.default_rng(4)
    shape = (24, 20, 28)
    f0 = None
    Ls = []
    for nt in (0, 1, 2):
        L = Lattice(*shape, periodic=(1, 0, 1), storage=storage, omega=1 / 0.8)
        mask = np.zeros(shape, np.uint8)
        mask[:, 0, :] = mask[:, -1, :] = 1
        L.set_mask(mask)
        L.set_body_force((1e-5, 0, 0))
        if f0 is None:
            L.init_equilibrium(rng.uniform(0.99, 1.01, shape), rng.uniform(-0.02, 0.02, shape + (3,)))
            f0 = L.get_populations()
        else:
            L.put_populations(f0)
        L.set_launch_parity(0, 128, False, nt)
        L.set_launch_parity(1, 192, 16, nt)
        assert L.launch(1)["nontemporal"] == nt and L.launch(0)["nontemporal"] == min(nt, 1)
        L.collide_and_stream(5)
        Ls.append(L)
    for L in Ls[1:]:
        assert np.array_equal(Ls[0].get_populations(), L.get_populations())


def test_coupled_tuner_layouts_give_the_same_trajectory(device):
    from biblbm_amd import tuning
    a, b = _box(seed=3), _box(seed=3)
    rep = tuning.tune_coupled(a.ibm, 4)
    n = rep["steps_run"]
    assert n == 2 * 4 * len(rep["trials"])
    b.ibm.coupled_steps(n, use_graph=False)
    assert np.array_equal(a.ibm.positions(), b.ibm.positions())
    assert np.array_equal(a.lattice.get_populations(), b.lattice.get_populations())
    for w in (a, b):
        w.ibm.close()
        w.lattice.close()


def test_tuning_table_save_load(device, tmp_path):
    from biblbm_amd import tuning
    from biblbm_amd.lattice import Lattice
    L = Lattice(40, 36, 44, periodic=(1, 1, 1), storage="f32", omega=1.0)
    rep = tuning.tune_collide(L)
    path = tmp_path / "tune.jsonl"
    tuning.save(str(path))
    lines = path.read_text().strip().splitlines()
    assert any('"nx": 40, "ny": 36, "nz": 44, "storage": 1' in ln for ln in lines)
    assert tuning.load(str(path))
    L2 = Lattice(40, 36, 44, periodic=(1, 1, 1), storage="f32", omega=1.0)
    assert L2.launch(0) == {k: rep["best"]["even"][k] for k in ("block", "xcd_remap", "nontemporal")}
    assert L2.launch(1)["block"] == rep["best"]["odd"]["block"]
    bad = tmp_path / "bad.jsonl"
    bad.write_text('{"kind": "collide", "device": "x", "nx": 1}\n')
    from biblbm_amd import capi
    with pytest.raises(capi.BlError):
        tuning.load(str(bad))


@pytest.mark.parametrize("order", [0, 1, 2, 3, 4, 16, 64, 1024])
def test_block_orders_are_bitwise(device, order):
    """Every block order (hardware order, one range per XCD, chunks of k
    blocks over the XCDs, whole z-rows per XCD) gives the same populations,
    on a lattice whose block count is not a multiple of the chunk round
    (tail blocks) and whose row count is not a multiple of 8."""
    from biblbm_amd.lattice import Lattice
    rng = np.random.default_rng(7)
    shape = (30, 26, 22)  # 17160 cells: 67.03 blocks of 256
    ref = None
    f0 = O_start = None
    L0 = Lattice(*shape, periodic=(1, 1, 0), storage="f64", omega=1 / 0.7)
    L0.init_equilibrium(rng.uniform(0.99, 1.01, shape), rng.uniform(-0.02, 0.02, shape + (3,)))
    f0 = L0.get_populations()
    L0.set_body_force((1e-5, 2e-6, 0))
    L0.collide_and_stream(4)
    ref = L0.get_populations()
    L = Lattice(*shape, periodic=(1, 1, 0), storage="f64", omega=1 / 0.7)
    L.put_populations(f0)
    L.set_body_force((1e-5, 2e-6, 0))
    for p in (0, 1):
        L.set_launch_parity(p, 64 if p else 128, order, False)
    L.collide_and_stream(4)
    assert np.array_equal(L.get_populations(), ref)


@pytest.mark.parametrize("shape,periodic,block,storage", [
    ((9, 5, 150), (1, 1, 1), 64, "f64"), ((9, 5, 150), (1, 0, 0), 128, "f64"), ((7, 3, 256), (0, 1, 1), 256, "f64"),
    ((4, 6, 64), (1, 1, 1), 192, "f64"), ((5, 4, 70), (0, 0, 0), 64, "f32"), ((6, 5, 90), (1, 1, 0), 128, "f16s")])
def test_row_orders_are_bitwise(device, shape, periodic, block, storage):
    """BL_ORDER_ROWS (rows longer than a block: several segments per row, the
    last one partial; a row count that is not a multiple of 8; reflection at
    bounded x, y and z faces), both parities, with every nontemporal level,
    against the hardware order."""
    from biblbm_amd.lattice import Lattice
    n = int(np.prod(shape))
    rho = 1 + 0.01 * np.sin(np.arange(n)).reshape(shape)
    u = 0.02 * np.cos(np.arange(3 * n)).reshape(shape + (3,))
    mask = np.zeros(shape, np.uint8)
    mask[1, 1, 1] = 1  # one bounce-back cell
    runs = []
    for order, nt in ((0, 0), (3, 0), (3, 1), (3, 2), (16, 2)):
        L = Lattice(*shape, periodic=periodic, storage=storage, omega=1 / 0.8)
        L.set_mask(mask)
        L.init_equilibrium(rho, u)
        L.set_body_force((1e-5, -2e-6, 3e-6))
        for p in (0, 1):
            L.set_launch_parity(p, block, order, nt)
            assert L.launch(p) == dict(block=block, xcd_remap=order, nontemporal=min(nt, 1) if p == 0 else nt)
        L.collide_and_stream(5)
        runs.append(L.get_populations())
        L.close()
    for r in runs[1:]:
        assert np.array_equal(runs[0], r)


@pytest.mark.parametrize("shape,periodic,storage,lanes,order,nt", [
    ((6, 5, 64), (1, 1, 1), "f64", 64, 0, 1), ((7, 4, 128), (0, 1, 1), "f64", 64, 64, 0),
    ((4, 4, 384), (1, 1, 1), "f64", 128, 64, 1), ((5, 6, 192), (1, 0, 1), "f32", 64, 1, 1),
    ((4, 5, 64), (0, 0, 1), "f16s", 64, 4, 0), ((3, 4, 512), (1, 1, 1), "f32", 256, 16, 2),
    ((5, 4, 384), (0, 1, 1), "f16s", 192, 0, 1), ((4, 3, 256), (1, 1, 1), "f64", 256, 0, 0),
    ((3, 3, 64), (1, 0, 1), "f16s", 64, 0, 0), ((3, 5, 128), (0, 1, 1), "f32", 128, 4, 1)])
def test_z_segment_kernel_is_bitwise(device, shape, periodic, storage, lanes, order, nt):
    """The odd step's z-segment kernel (block -B: one workgroup per B cells
    of a z-row, every access z-aligned, the z shift through LDS; one or
    several segments per row, dealt in several orders) against the per-cell
    kernel: reflection at bounded x and y faces, bounce-back cells, a body
    force, both parities over five steps."""
    from biblbm_amd.lattice import Lattice
    n = int(np.prod(shape))
    rho = 1 + 0.01 * np.sin(np.arange(n)).reshape(shape)
    u = 0.02 * np.cos(np.arange(3 * n)).reshape(shape + (3,))
    mask = np.zeros(shape, np.uint8)
    mask[1, 1, 1] = 1
    mask[0, shape[1] - 1, shape[2] - 1] = 1
    mask[2, 0, lanes - 1] = 1
    runs = []
    for block in (lanes, -lanes):
        L = Lattice(*shape, periodic=periodic, storage=storage, omega=1 / 0.8)
        L.set_mask(mask)
        L.init_equilibrium(rho, u)
        L.set_body_force((1e-5, -2e-6, 3e-6))
        L.set_launch_parity(1, block, order, nt)
        assert L.launch(1) == dict(block=block, xcd_remap=order, nontemporal=nt)
        L.collide_and_stream(5)
        runs.append(L.get_populations())
        L.close()
    assert np.array_equal(runs[0], runs[1])


def test_z_segment_kernel_falls_back_where_it_does_not_apply(device):
    """A negative block on a lattice the z-segment kernel cannot take (nz
    not a multiple of the segment; z bounded) steps as the positive block
    does; BL_ORDER_ROWS is refused with it."""
    from biblbm_amd import capi
    from biblbm_amd.lattice import Lattice
    for shape, periodic in (((6, 5, 40), (1, 1, 1)), ((6, 5, 64), (1, 1, 0)), ((4, 4, 96), (1, 1, 1))):
        n = int(np.prod(shape))
        rho = 1 + 0.01 * np.sin(np.arange(n)).reshape(shape)
        runs = []
        for block in (64 if shape[2] != 96 else 192, -64 if shape[2] != 96 else -192):
            L = Lattice(*shape, periodic=periodic, storage="f64", omega=1 / 0.8)
            L.init_equilibrium(rho, None)
            L.set_body_force((1e-5, 0, 0))
            for p in (0, 1):
                L.set_launch_parity(p, block, 0, 0)
            L.collide_and_stream(4)
            runs.append(L.get_populations())
            if block < 0:
                with pytest.raises(capi.BlError):
                    L.set_launch_parity(1, block, 3, 0)
            L.close()
        assert np.array_equal(runs[0], runs[1])


@pytest.mark.parametrize("storage,side", [("f64", 64), ("f16s", 128)])
def test_z_segment_kernel_in_coupled_steps(device, storage, side):
    """Coupled steps of a shear box (velocity walls, tile-gated force field;
    one and two segments per row) with the z-segment odd step, the velocity
    staging carried by the collision and by a pass of its own, eagerly and
    from captured graphs, against the per-cell odd step."""
    from biblbm_amd import workloads
    ws = []
    for block in (192, -64):
        for fused in (1, 0):
            w = workloads.shear_box(storage, side, 0.9, 0.3, G=0.004, seed=5, radius=9.0)
            w.lattice.set_launch_parity(1, block, 64, 1)
            w.ibm.set_fused_staging(fused)
            w.ibm.coupled_steps(3, use_graph=False)
            w.ibm.coupled_steps(6, use_graph=True)
            ws.append(w)
    for w in ws[1:]:
        assert np.array_equal(ws[0].lattice.get_populations(), w.lattice.get_populations())
        assert np.array_equal(ws[0].ibm.positions(), w.ibm.positions())
    for w in ws:
        w.ibm.close()
        w.lattice.close()


def _small_suspension(storage="f64", seed=3):
    from biblbm_amd import workloads
    return workloads.suspension(storage, nx=64, ny=40, nz=48, n642=2, n2562=1, n162=2, cell_radius=8.0, seed=seed)


ORDER_KEEPING = [dict(bin_block=64), dict(bin_block=128), dict(elem_block=64), dict(elem_block=128),
                 dict(interp_block=64), dict(interp_block=128), dict(spread_small=1), dict(spread_small=8),
                 dict(spread_big=16), dict(spread_big=128), dict(xcd=2), dict(xcd=4), dict(xcd=6)]


@pytest.mark.parametrize("fused,layered", [(0, 1), (1, 0), (0, 0)])
def test_ibm_launch_parameters_are_bitwise(device, fused, layered):
    """Every immersed-boundary launch parameter except the gather lane count
    leaves the trajectory bitwise the same, for each step layout (paired
    stages on one stream, staging carried by the collision, forked)."""
    ref = _small_suspension()
    ref.ibm.set_fused_staging(fused)
    ref.ibm.set_layered(layered)
    ref.ibm.coupled_steps(6, use_graph=True)
    x_ref, f_ref = ref.ibm.positions(), ref.lattice.get_populations()
    ref.ibm.close()
    ref.lattice.close()
    for params in ORDER_KEEPING:
        W = _small_suspension()
        W.ibm.set_fused_staging(fused)
        W.ibm.set_layered(layered)
        W.ibm.set_launch(**params)
        W.ibm.coupled_steps(2, use_graph=False)
        W.ibm.coupled_steps(4, use_graph=True)
        assert np.array_equal(W.ibm.positions(), x_ref), params
        assert np.array_equal(W.lattice.get_populations(), f_ref), params
        W.ibm.close()
        W.lattice.close()


@pytest.mark.parametrize("lanes", [2, 8])
def test_gather_lane_counts_agree_to_rounding(device, lanes):
    """2 or 8 lanes per vertex in the force gather change only the order of
    the per-vertex sums: FP64 trajectories agree with the 4-lane default to
    1e-12 relative, and each lane count is bitwise repeatable."""
    runs = []
    for gl in (4, lanes, lanes):
        W = _small_suspension()
        W.ibm.set_launch(gather_lanes=gl)
        p0 = W.ibm.positions()
        W.ibm.coupled_steps(6, use_graph=True)
        runs.append((W.ibm.positions(), p0))
        W.ibm.close()
        W.lattice.close()
    (x4, p0), (xa, _), (xb, _) = runs
    assert np.array_equal(xa, xb)
    assert np.linalg.norm(xa - x4) <= 1e-12 * np.linalg.norm(x4 - p0)


def test_ibm_launch_rejects_values_outside_the_sets(device):
    from biblbm_amd import capi
    W = _small_suspension()
    before = W.ibm.launch()
    for bad in (dict(bin_block=512), dict(gather_lanes=3), dict(spread_small=16), dict(spread_big=8),
                dict(interp_block=32), dict(elem_block=0), dict(xcd=3), dict(xcd=8)):
        with pytest.raises(capi.BlError):
            W.ibm.set_launch(**bad)
        assert W.ibm.launch() == before
    W.ibm.close()
    W.lattice.close()


def test_ibm_tuner_persists_and_applies(device, tmp_path):
    """bl_tune_ibm picks launch parameters from its trials (the report
    counts the steps it ran), records them under the boundary's key, and a
    saved table applied to a fresh boundary of the same workload sets them."""
    from biblbm_amd import tuning
    W = _small_suspension("f32")
    s0 = W.lattice.steps
    rep = W.ibm.tune_launch(2)
    best = rep["best"]
    assert len(rep["trials"]) == 34  # two passes over every other value of each parameter
    assert rep["steps_run"] == 2 * 2 * (1 + len(rep["trials"]))
    assert W.ibm.launch() == {k: best[k] for k in W.ibm.LAUNCH_KEYS}
    assert W.lattice.steps - s0 == rep["steps_run"]
    path = tmp_path / "tune.jsonl"
    tuning.save(str(path))
    assert '"ibm_launch": [' in path.read_text()
    W.ibm.close()
    W.lattice.close()
    assert tuning.load(str(path))
    W2 = _small_suspension("f32")
    tuning.apply_coupled(W2.ibm)
    assert W2.ibm.launch() == {k: best[k] for k in W2.ibm.LAUNCH_KEYS}
    W2.ibm.close()
    W2.lattice.close()
