#!/usr/bin/env python3
"""Benchmark: coupled D3Q19 + immersed-boundary time steps on one MI355X.

Headline workload (the largest single-GPU configuration): the precision
configuration, a 384^3 periodic box seeded with a Taylor-Green field
(u0 0.03, tau 0.8) holding 16 Skalak capsules (642-vertex icospheres, radius
8, Ca 0.1) at seeded non-overlapping positions, Peskin 4-point kernel,
populations stored in FP64.  A step is one full coupled step (stencil
binning, velocity staging, interpolation, membrane forces, spreading,
advection, collide-and-stream), replayed from captured hipGraphs that are
instantiated before the timed region; inputs are resident in HBM before it
starts.

value = lattice cells x steps / seconds / 1e6 (MLUPS).  Extra objects on the
same JSON line:
  roofline      collide-stream kernel (the dominant kernel) on the headline
                lattice: algorithmic bytes per launch / average launch time
                from HIP events on the lattice stream, against 8 TB/s; with
                the per-parity times, the tuned launch shapes and the
                measured device-copy bandwidth at 4 GiB
  cpu_baseline  the reference's own lattice passes of a coupled step
                (compiled from its sources, serial) on the headline lattice,
                two steps; beside it (port_coupled_step) the FP64 C
                restatement (oracle) running the whole coupled step on this
                host's cores, and the GPU's vertex positions after the same
                steps against it
  precision     the headline workload in FP32 and FP16-shifted storage from
                the same start: speed and velocity / position / deformation
                index errors against the FP64 run
  capsule_in_shear  the 64^3 shear box with one capsule (FP64)
  rbc_in_pipe   pipe R 32 x L 257, one red cell, FP64 and FP32, and with the
                reference prog's default Guo curved wall (second_order);
                its populations (175 MB FP64) fit the 256 MB last-level
                cache, so its roofline figure is cache-assisted
  suspension    the 400-cell channel: FP64 / FP32 / FP16-shifted speed,
                velocity error against the FP64 run, per-stage IBM times
  interactions  the suspension with Morse wall and particle-particle forces
  ensemble      64 shear boxes of uneven sizes advanced by one hipGraph
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLUPS + %HBM roofline, D3Q19+IBM, 1 GPU; FP16-store vs FP32 at fixed error"
HBM_PEAK_GBS = 8000.0
BYTES = {"f64": 8, "f32": 4, "f16s": 2}


def progress(msg):
    """One line on stderr per finished leg (the JSON line stays alone on stdout)."""
    sys.stderr.write("bench: %s\n" % msg)
    sys.stderr.flush()


def collide_bytes_per_launch(cells, storage, has_mask, active_tiles, ntiles):
    """Algorithmic bytes of one collide-and-stream launch: 2 x 19 populations
    per cell, the mask byte if the lattice has one, the force field of the
    IBM-touched tiles (3 values per cell) and one flag byte per tile."""
    s = BYTES[storage]
    sF = 8 if storage == "f64" else 4
    b = cells * 2 * 19 * s
    if has_mask:
        b += cells
    b += active_tiles * 64 * 3 * sF + ntiles
    return b


def read_traffic(workload, storage):
    """HBM bytes per collide-stream launch measured with rocprofv3 counters
    (FETCH_SIZE + WRITE_SIZE, calibrated against a copy kernel), as recorded
    in profiles/; None if that profile is absent."""
    path = os.path.join(ROOT, "profiles", "collide_stream_%s_%s_counters.md" % (workload, storage))
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def ntiles_of(L):
    nx, ny, nz = L.shape
    return ((nx + 3) // 4) * ((ny + 3) // 4) * ((nz + 3) // 4)


def time_coupled(W, steps, warmup):
    """warmup steps, then the graphs of the timed run are captured
    (bl_coupled_prepare) and `steps` coupled steps are timed between two
    device synchronisations."""
    from biblbm_amd import capi
    W.ibm.coupled_steps(warmup, use_graph=True)
    capi.call("bl_sync")
    W.ibm.prepare(steps)
    capi.call("bl_sync")
    t0 = time.perf_counter()
    W.ibm.coupled_steps(steps, use_graph=True)
    capi.call("bl_sync")
    return time.perf_counter() - t0


def roofline(W, storage, has_mask, n=20):
    """The collide kernel alone on the workload's lattice (the force field
    and flags of the last step in place): n launches over both parities
    timed with HIP events on the lattice stream, then each parity alone."""
    L = W.lattice
    st = W.ibm.tile_stats()
    ms = L.time_collide(n) / n
    par = [0.0, 0.0]
    reps = 4
    for _ in range(reps):
        for _ in range(2):
            par[L.steps & 1] += L.time_collide(1) / reps
    nb = collide_bytes_per_launch(W.cells, storage, has_mask, st["active_tiles"], ntiles_of(L))
    achieved = nb / (ms * 1e-3) / 1e9
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=read_traffic(W.name, storage),
                kernel="bl::k_collide_stream" + (" (even steps) / bl::k_collide_zseg (odd steps)"
                                                 if L.launch(1)["block"] < 0 else ""),
                launch_ms=round(ms, 5), algorithmic_bytes=nb,
                bytes_per_cell=2 * 19 * BYTES[storage] + (1 if has_mask else 0),
                active_tiles=st["active_tiles"],
                by_parity={p: dict(launch_ms=round(par[k], 5), frac=round(nb / (par[k] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                   **L.launch(k)) for k, p in ((0, "even"), (1, "odd"))})


def copy_bandwidth(nbytes=1 << 32, n_iter=10):
    """Device-to-device copy of nbytes (16-byte elements), read + write bytes
    per second: the bandwidth a streaming kernel can reach on this device."""
    import ctypes
    from biblbm_amd import capi
    ms = ctypes.c_double()
    capi.call("bl_device_copy", nbytes, n_iter, ctypes.byref(ms))
    return round(2.0 * nbytes * n_iter / (ms.value * 1e-3) / 1e9, 1)


def oracle_particles(ibm):
    from oracle import oracle_py as O
    pos = ibm.positions()
    out, k = [], 0
    for sid, model, params, bug in ibm.specs:
        rest, tri = ibm.meshes[sid]
        out.append(O.Particle(rest, tri, model, params, pos[k:k + len(rest)], bug))
        k += len(rest)
    return out


REF_TIME = os.path.join(ROOT, "oracle", "_ref", "ref_time")


def reference_baseline(steps=2, timeout_s=240):
    """The reference's own lattice passes of one coupled step, compiled from
    its sources (oracle/_ref/ref_time: computeVelocity, setExternalVector,
    collideAndStream with GuoExternalForceVOFDynamics; serial, one core) on
    the headline lattice (384^3 Taylor-Green, tau 0.8), `steps` steps.
    None when the binary is absent."""
    import subprocess
    if not os.path.exists(REF_TIME):
        return None
    r = subprocess.run([REF_TIME, "384", str(steps), "0.8", "0.03"], capture_output=True, text=True,
                       timeout=timeout_s)
    if r.returncode != 0:
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_baseline(steps, budget_s):
    """The oracle (FP64 C restatement, OpenMP in the lattice passes, IBM
    serial) on the headline workload from its start: same lattice, fields,
    meshes, parameters and positions; then the GPU runs the same number of
    steps from a fresh start and its vertex positions are compared."""
    import numpy as np
    from biblbm_amd import workloads
    from oracle import oracle_py as O
    G = workloads.taylor_green("f64")
    n = G.params["n"]
    rho, u = workloads.taylor_green_fields(n, G.params["u0"])
    f = O.lbm_init_equilibrium(rho, u)
    del rho, u
    parts = oracle_particles(G.ibm)
    shape = (n, n, n)
    force = np.zeros(shape + (3,))
    zero = np.zeros(shape)
    mask = np.zeros(shape, np.uint8)
    om = 1 / G.params["tau"]
    done = 0
    t0 = time.perf_counter()
    while done < steps:
        f, force, _ = O.coupled_step(f, force, zero, mask, (1, 1, 1), om, om, (0, 0, 0), parts, "peskin")
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    del f, force
    G.ibm.coupled_steps(done, use_graph=False)
    gpos = G.ibm.positions()
    ref = np.concatenate([p.pos for p in parts])
    G.ibm.close()
    G.lattice.close()
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=round(n ** 3 * done / dt / 1e6, 3), unit="MLUPS", cores=cores, kind="port",
                sample="%d coupled steps of the 384^3 Taylor-Green box with 16 capsules (56 623 104 cells, "
                       "10 272 vertices), FP64 oracle, OpenMP in the lattice passes" % done,
                gpu_vs_cpu_position_rel_l2=float(np.linalg.norm(gpos - ref) / np.linalg.norm(ref)))


def precision_leg(steps, checkpoints=(100, 200)):
    """The headline workload in FP32 and FP16-shifted storage (FP32
    collision, FP64 rest geometry and conservation sums) from the same start
    as the FP64 run: errors against it after each checkpoint step count and
    after `steps` (velocity field rel-L2, vertex positions rel-L2, largest
    per-capsule change of the Taylor deformation index), speed from the
    steps after the first checkpoint; each storage's collide launch shapes
    are tuned first (the tuner restores the state)."""
    import numpy as np
    from biblbm_amd import capi, workloads
    from biblbm_amd.ibm import taylor_deformation
    marks = sorted(set([c for c in checkpoints if c < steps] + [steps]))
    out = {}
    ref = {}
    from biblbm_amd import tuning
    for storage in ("f64", "f32", "f16s"):
        W = workloads.taylor_green(storage)
        row = {"collide_launch": tuning.tune_collide(W.lattice)["best"]}
        done, t_timed, n_timed = 0, 0.0, 0
        for m in marks:
            W.ibm.prepare(m - done)
            capi.call("bl_sync")
            t0 = time.perf_counter()
            W.ibm.coupled_steps(m - done, use_graph=True)
            capi.call("bl_sync")
            if done > 0:
                t_timed += time.perf_counter() - t0
                n_timed += m - done
            done = m
            _, u = W.lattice.moments()
            pos = W.ibm.positions()
            D = np.array([taylor_deformation(W.ibm.observables(p)["inertia"]) for p in range(16)])
            if storage == "f64":
                ref[m] = (u, pos, D)
                row["deformation_index_at_%d" % m] = dict(min=float(D.min()), max=float(D.max()))
            else:
                r = ref[m]
                row["at_%d_steps" % m] = dict(
                    velocity_rel_l2_vs_f64=float(np.linalg.norm(u - r[0]) / np.linalg.norm(r[0])),
                    position_rel_l2_vs_f64=float(np.linalg.norm(pos - r[1]) / np.linalg.norm(r[1])),
                    deformation_index_max_abs_diff_vs_f64=float(np.abs(D - r[2]).max()))
            del u
        row["mlups"] = round(W.cells * n_timed / t_timed / 1e6, 1)
        row["ms_per_step"] = round(t_timed / n_timed * 1e3, 4)
        out[storage] = row
        W.ibm.close()
        W.lattice.close()
        del W
    ref.clear()
    out["fp16s_over_fp32"] = round(out["f16s"]["mlups"] / out["f32"]["mlups"], 3)
    out["workload"] = ("taylor_green 384^3 periodic, u0 0.03, tau 0.8, 16 Skalak capsules (642 vertices, "
                       "radius 8, Ca 0.1), Peskin, %s steps from the same start" % "/".join(map(str, marks)))
    return out


def suspension_leg(steps, warmup, horizons=(500, 1000, 2000), tune_steps=4):
    """The 400-cell channel in FP64, FP32 and FP16-shifted storage from the
    same start (the immersed boundary in FP64 with FP64 populations, FP32
    otherwise): the collide launch shapes (bl_tune_collide, which restores
    the state) and the immersed-boundary launch parameters are tuned first
    (bl_tune_ibm, the same number of real steps in every storage), then the
    speed over `steps` after `warmup`, then the runs continue to each
    horizon (steps from the start) and the velocity field and vertex
    positions are compared with the FP64 run's there."""
    import numpy as np
    from biblbm_amd import tuning, workloads
    out = {}
    ref = {}
    for storage in ("f64", "f32", "f16s"):
        W = workloads.suspension(storage)
        collide_tuned = tuning.tune_collide(W.lattice)["best"]
        tuned = tuning.tune_ibm(W.ibm, tune_steps)["best"]
        dt = time_coupled(W, steps, warmup)
        done = W.lattice.steps
        errs = {}
        for h in horizons:
            if h < done:
                raise RuntimeError("suspension leg: horizon %d is before the timed steps end (%d)" % (h, done))
            if h > done:
                W.ibm.coupled_steps(h - done, use_graph=True)
                done = h
            _, u = W.lattice.moments()
            x = W.ibm.positions()
            if storage == "f64":
                ref[h] = (u, x)
            else:
                errs["at_%d_steps" % h] = dict(
                    velocity_rel_l2_vs_f64=float(np.linalg.norm(u - ref[h][0]) / np.linalg.norm(ref[h][0])),
                    position_rel_l2_vs_f64=float(np.linalg.norm(x - ref[h][1]) / np.linalg.norm(ref[h][1])))
            del u, x
        W.ibm.step()
        stages = {k: round(v, 5) for k, v in W.ibm.stage_times().items()}
        row = dict(mlups=round(W.cells * steps / dt / 1e6, 1), ms_per_step=round(dt / steps * 1e3, 4),
                   ibm_stage_ms=stages, ibm_launch=tuned, collide_launch=collide_tuned)
        if storage != "f64":
            row["velocity_rel_l2_vs_f64"] = errs["at_%d_steps" % horizons[0]]["velocity_rel_l2_vs_f64"]
            row["errors"] = errs
        out[storage] = row
        nv, vf = W.vertices, W.params["volume_fraction"]
        W.ibm.close()
        W.lattice.close()
        del W
    ref.clear()
    out["fp16s_over_fp32"] = round(out["f16s"]["mlups"] / out["f32"]["mlups"], 3)
    out["workload"] = ("suspension 512x128x128 channel, 400 cells (300x642, 80x2562, 20x162 vertices), "
                       "%.1f%% volume fraction, %d vertices, %d steps timed after the launch tuner's steps and %d "
                       "warmup, errors against the FP64 run up to %d steps" % (100 * vf, nv, steps, warmup,
                                                                             max(horizons)))
    return out


def interactions_leg(steps, warmup, f32_ms=None):
    """The 400-cell suspension (FP32 storage) with the rbc_in_pipe prog's
    interactions: Morse wall decorator on the channel walls, Morse
    particle-particle force through the collision handler."""
    from biblbm_amd import capi, workloads
    from biblbm_amd.potentials import MorsePotential
    W = workloads.suspension("f32", interactions=True)
    dt = time_coupled(W, steps, warmup)
    u = MorsePotential(1.0, 1.0, 1e-3)
    W.ibm.compute_collision_forces(u)
    capi.call("bl_sync")
    t0 = time.perf_counter()
    for _ in range(20):
        W.ibm.compute_collision_forces(u)
    capi.call("bl_sync")
    pass_ms = (time.perf_counter() - t0) / 20 * 1e3
    out = dict(mlups=round(W.cells * steps / dt / 1e6, 1), ms_per_step=round(dt / steps * 1e3, 4),
               collision_pass_ms=round(pass_ms, 4), vertices=W.vertices,
               workload="suspension 512x128x128 FP32 + Morse wall (2, 2, 1e-4) + Morse pair force (1, 1, 1e-3)")
    if f32_ms:
        out["overhead_vs_f32_suspension"] = round(dt / steps * 1e3 / f32_ms - 1.0, 4)
    W.ibm.close()
    W.lattice.close()
    return out


def ensemble_oracle_check(boxes=(0, 1), steps=20):
    """Boxes of the ensemble configuration (the first two draws), stepped as
    a batched ensemble of their own, against the FP64 oracle's composed
    steps from the same start: vertex positions rel-L2 and the Taylor
    deformation index of each box."""
    import numpy as np
    from biblbm_amd import meshes, workloads
    from biblbm_amd.ensemble import Ensemble
    from biblbm_amd.ibm import taylor_deformation
    from oracle import oracle_py as O
    params = workloads.shear_ensemble_params()
    ws = [workloads.shear_box("f64", s, tau, Ca, seed=bs) for s, tau, Ca, bs in (params[k] for k in boxes)]
    starts = [(w.lattice.get_populations(), w.ibm.positions()) for w in ws]
    ens = Ensemble([(w.lattice, w.ibm) for w in ws])
    ens.steps(steps)
    out = []
    for w, (f, pos0) in zip(ws, starts):
        shape = w.lattice.shape
        mask = np.zeros(shape, np.uint8)
        mask[:, 0, :] = mask[:, -1, :] = 3
        u_w = w.params["wall_u"]
        wall_u = np.array([-u_w, 0, 0, u_w, 0, 0])
        R, G = w.params["radius"], w.params["G"]
        v, t = meshes.biconcave(3, R)
        p = O.Particle(v, t, 3, workloads.skalak_rbc_params(G, 10.0, R, meshes.signed_volume(v, t), 0.02 * G * R * R),
                       pos0)
        force = np.zeros(shape + (3,))
        om = 1 / w.params["tau"]
        for _ in range(steps):
            f, force, _ = O.coupled_step(f, force, np.zeros(shape), mask, (1, 0, 1), om, om, (0, 0, 0), [p], "peskin",
                                         1, wall_u)
        gpos = w.ibm.positions()
        D_gpu = taylor_deformation(w.ibm.observables(0)["inertia"])
        D_ref = taylor_deformation(O.moment_of_inertia(t, p.pos))
        out.append(dict(side=w.params["s"], steps=steps,
                        position_rel_l2=float(np.linalg.norm(gpos - p.pos) / np.linalg.norm(p.pos - pos0)),
                        deformation_index=D_gpu, deformation_index_abs_diff_vs_oracle=float(abs(D_gpu - D_ref))))
    ens.close()
    for w in ws:
        w.ibm.close()
        w.lattice.close()
    return out


def ensemble_leg(steps, warmup, n_boxes=64):
    """64 independent shear boxes (sides from {32..64}^3, one 642-vertex cell
    each): every stage of a step one launch over all boxes (batched), two
    steps per hipGraph launch, the batched collide's workgroup order from
    the ensemble tuner; against the same boxes stepped one after another
    (each with its own captured graph), against the collide-and-stream rate
    of one lattice of the same total cells, and two boxes against the
    oracle."""
    import numpy as np
    from biblbm_amd import capi, workloads
    from biblbm_amd.ensemble import Ensemble
    from biblbm_amd.ibm import taylor_deformation
    ws = workloads.shear_ensemble("f64", n_boxes=n_boxes)
    cells = sum(w.cells for w in ws)
    ens = Ensemble([(w.lattice, w.ibm) for w in ws])
    ens.steps(warmup)
    tuned = ens.tune(4)
    ms = ens.time_steps(steps)
    batched = ens.batched
    for w in ws:
        w.ibm.coupled_steps(warmup, use_graph=True)
        w.ibm.prepare(steps)
    capi.call("bl_sync")
    t0 = time.perf_counter()
    for w in ws:
        w.ibm.coupled_steps(steps, use_graph=True)
    capi.call("bl_sync")
    loop_s = time.perf_counter() - t0
    D = np.array([taylor_deformation(w.ibm.observables(0)["inertia"]) for w in ws])
    out = dict(mlups=round(cells * steps / (ms * 1e-3) / 1e6, 1), ms_per_step=round(ms / steps, 4),
               batched=batched, tuned_collide_order=tuned["best"],
               per_box_loop_mlups=round(cells * steps / loop_s / 1e6, 1),
               deformation_index=dict(min=float(D.min()), median=float(np.median(D)), max=float(D.max())),
               workload="%d shear boxes, %d cells, sides %s, one 642-vertex biconcave cell each, FP64, "
                        "steps %d after %d warmup and the tuner's steps" %
                        (n_boxes, cells, sorted({w.params["s"] for w in ws}), steps, warmup))
    ens.close()
    for w in ws:
        w.ibm.close()
        w.lattice.close()
    # the single-lattice collide rate at equal total cells: one shear box of
    # that many cells (the same walls and storage), its collide launch
    # shapes tuned, the collide-and-stream alone over both parities
    from biblbm_amd import tuning
    side = int(round(cells ** (1.0 / 3.0)))
    W1 = workloads.shear_box("f64", side, 0.9, 0.3, G=0.004, seed=0, radius=0.2 * side)
    W1.ibm.coupled_steps(2, use_graph=False)
    tuning.tune_collide(W1.lattice)
    n1 = 20
    ms1 = W1.lattice.time_collide(n1) / n1
    single = W1.cells / (ms1 * 1e-3) / 1e6
    out["single_lattice_collide"] = dict(side=side, cells=W1.cells, ms=round(ms1, 4), mlups=round(single, 1))
    out["vs_single_lattice_collide"] = round(out["mlups"] / single, 3)
    W1.ibm.close()
    W1.lattice.close()
    out["vs_oracle"] = ensemble_oracle_check()
    return out


def pipe_leg(steps, warmup):
    from biblbm_amd import workloads
    out = {}
    for storage in ("f64", "f32"):
        W = workloads.rbc_in_pipe(storage)
        dt = time_coupled(W, steps, warmup)
        rl = roofline(W, storage, True)
        rl["note"] = "populations of this lattice fit the 256 MB last-level cache: cache-assisted figure"
        # the same run continued with the inert bounce-back cells beyond the
        # wall left alone (bl_lattice_set_skip_inert; fluid results bitwise)
        n_inert = W.lattice.set_skip_inert(1)
        dts = time_coupled(W, steps, warmup)
        W.lattice.set_skip_inert(0)
        out[storage] = dict(mlups=round(W.cells * steps / dt / 1e6, 2), ms_per_step=round(dt / steps * 1e3, 5),
                            fluid_mlups=round(W.fluid_cells * steps / dt / 1e6, 2), roofline=rl,
                            skip_inert=dict(inert_cells=n_inert, ms_per_step=round(dts / steps * 1e3, 5),
                                            mlups=round(W.cells * steps / dts / 1e6, 2),
                                            fluid_mlups=round(W.fluid_cells * steps / dts / 1e6, 2)))
        W.ibm.close()
        W.lattice.close()
    W2 = workloads.rbc_in_pipe("f64", second_order=True)
    n2 = max(2, steps // 2)
    dt2 = time_coupled(W2, n2, max(2, warmup // 2))
    out["second_order"] = dict(mlups=round(W2.cells * n2 / dt2 / 1e6, 2), ms_per_step=round(dt2 / n2 * 1e3, 5),
                               wall_nodes=W2.params["wall_nodes"], fluid_cells=W2.fluid_cells)
    W2.ibm.close()
    W2.guo.close()
    W2.lattice.close()
    out["workload"] = ("rbc_in_pipe: pipe R=32 L=257 (257x67x67 cells, bounce-back; second_order: Guo curved "
                       "wall, no dynamics beyond it), tau 0.9, body force for u_c 0.02, one 2562-vertex "
                       "Evans-Fung red cell R=12 (Skalak + bending + area + volume), Peskin")
    return out


def capsule_leg(steps, warmup):
    from biblbm_amd import workloads
    W = workloads.capsule_in_shear("f64")
    dt = time_coupled(W, steps, warmup)
    out = dict(mlups=round(W.cells * steps / dt / 1e6, 2), ms_per_step=round(dt / steps * 1e3, 5),
               workload="capsule_in_shear 64^3, velocity walls +-0.01, tau 1, Skalak capsule 642 vertices R=8, "
                        "Ca 0.1, Peskin, FP64")
    W.ibm.close()
    W.lattice.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--storage", default="f64", choices=["f64", "f32", "f16s"])
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--cpu-budget", type=float, default=25.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the launch shapes the lattice starts with (from the loaded tuning table, if it has "
                         "this size) instead of tuning them here")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        sys.exit("bench.py: this build runs on one GPU only (the path's multi-GPU sharding is a design, "
                 "not an implementation); run with --gpus 1")

    from biblbm_amd import capi, tuning, workloads

    capi.call("bl_init", 0)
    tuning.load()
    W = workloads.taylor_green(args.storage)
    if args.no_tune:
        tuned = {"even": W.lattice.launch(0), "odd": W.lattice.launch(1), "source": "as created"}
    else:
        tuned = tuning.tune_collide(W.lattice)["best"]
    progress("headline workload built, collide launch shapes %s" % ("kept" if args.no_tune else "tuned"))
    dt = time_coupled(W, args.steps, args.warmup)
    value = W.cells * args.steps / dt / 1e6
    W.ibm.step()
    stages = {k: round(v, 5) for k, v in W.ibm.stage_times().items()}
    rl = roofline(W, args.storage, False)
    rl["tuned"] = tuned
    rl["device_copy_gbs_4gib"] = copy_bandwidth()
    rl["copy_frac"] = round(rl["device_copy_gbs_4gib"] / HBM_PEAK_GBS, 4)
    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": args.storage, "data": "synthetic",
        "config": {"workload": "precision: 384^3 periodic Taylor-Green box (u0 0.03, tau 0.8) with 16 Skalak "
                               "capsules (642-vertex icospheres, radius 8, Ca 0.1) at seeded positions, Peskin "
                               "kernel, hipGraph-replayed coupled steps",
                   "cells": W.cells, "vertices": W.vertices, "storage": args.storage, "parallelism": "single GPU"},
        "roofline": rl,
        "ibm_stage_ms": stages,
        "vertex_updates_per_s": round(W.vertices * args.steps / dt, 1),
    }
    W.ibm.close()
    W.lattice.close()
    del W
    progress("headline done: %.1f MLUPS" % value)
    if not args.no_extras:
        line["capsule_in_shear"] = capsule_leg(max(50, args.steps), 10)
        line["rbc_in_pipe"] = pipe_leg(max(100, args.steps), 20)
        progress("capsule_in_shear / rbc_in_pipe done")
        line["suspension"] = suspension_leg(max(50, args.steps // 4), 10)
        progress("suspension done")
        line["interactions"] = interactions_leg(max(20, args.steps // 10), 10,
                                                line["suspension"]["f32"]["ms_per_step"])
        progress("interactions done")
        line["ensemble"] = ensemble_leg(max(50, args.steps // 4), 10)
        progress("ensemble done")
        line["precision"] = precision_leg(max(500, args.steps), checkpoints=(100, 200, 300))
        progress("precision done")
    if not args.no_cpu_baseline:
        port = cpu_baseline(args.cpu_steps, args.cpu_budget)
        ref = reference_baseline()
        if ref is not None:
            line["cpu_baseline"] = dict(
                value=round(ref["mlups"], 4), unit="MLUPS", cores=1, kind="reference",
                sample="%d steps of the reference's own lattice passes of a coupled step (computeVelocity, "
                       "setExternalVector, collideAndStream with GuoExternalForceVOFDynamics; Palabos serial build "
                       "from the reference sources) on the 384^3 Taylor-Green lattice; its immersed-boundary "
                       "class needs MPI and is not built, so the 16 capsules' stages are not in this figure"
                       % ref["steps"],
                seconds=round(ref["seconds"], 3), setup_seconds=ref["setup_seconds"], port_coupled_step=port)
        else:
            line["cpu_baseline"] = port
        progress("cpu baseline done")
    tuning.save()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
