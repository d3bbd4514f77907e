#!/usr/bin/env python3
"""Benchmark: coupled D3Q19 + immersed-boundary time steps on one MI355X.

Default workload: rbc_in_pipe (pipe radius 32, length 257, bbox
257 x 67 x 67 = 1 153 673 lattice cells, bounce-back wall, body force for a
centreline velocity of 0.02, tau 0.9, one 2562-vertex Evans-Fung red cell
with Skalak + bending + area + volume forces, Peskin 4-point kernel),
populations stored in FP64.  A step is one full coupled step (stencil
binning, velocity staging, interpolation, membrane forces, spreading,
advection, collide-and-stream), replayed from a captured hipGraph; inputs are
resident in HBM before the timed region.

value = lattice cells x steps / seconds / 1e6 (MLUPS over all cells of the
bounding box).  Extra objects on the same JSON line:
  roofline      collide-stream kernel (the dominant kernel): algorithmic bytes
                per launch / average launch time from HIP events on the
                lattice stream, against 8 TB/s
  cpu_baseline  the FP64 C restatement (oracle) running the same coupled step
                on this host's cores, on a bounded number of steps
  fp32          the same workload with FP32 storage
  suspension    the 400-cell channel: FP16-shifted vs FP32 storage speed and
                velocity error against the FP64 run
  ensemble      64 shear boxes of uneven sizes advanced by one hipGraph
                launch, against the same boxes stepped one after another
  second_order  rbc_in_pipe with the reference prog's default wall: the Guo
                curved wall on the pipe surface, no dynamics beyond it
  interactions  the suspension with the rbc_in_pipe prog's Morse wall and
                particle-particle interactions, and one collision-handler
                pass alone
  precision     384^3 Taylor-Green box with 16 capsules in FP64, FP32 and
                FP16-shifted storage from the same start: speed, and the
                velocity, position and deformation-index errors of the
                reduced forms against the FP64 run
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


def collide_bytes_per_launch(cells, storage, has_mask, active_tiles, ntiles):
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


def time_coupled(W, steps, warmup):
    from biblbm_amd import capi
    W.ibm.coupled_steps(warmup, use_graph=True)
    capi.call("bl_sync")
    t0 = time.perf_counter()
    W.ibm.coupled_steps(steps, use_graph=True)
    capi.call("bl_sync")
    return time.perf_counter() - t0


def roofline(W, storage, n=40):
    L = W.lattice
    st = W.ibm.tile_stats()
    ms = L.time_collide(n) / n
    nb = collide_bytes_per_launch(W.cells, storage, True, st["active_tiles"], L_ntiles(L))
    achieved = nb / (ms * 1e-3) / 1e9
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=read_traffic(W.name, storage),
                kernel="bl::k_collide_stream", launch_ms=round(ms, 5), algorithmic_bytes=nb,
                active_tiles=st["active_tiles"])


def L_ntiles(L):
    nx, ny, nz = L.shape
    return ((nx + 3) // 4) * ((ny + 3) // 4) * ((nz + 3) // 4)


def cpu_baseline(steps, budget_s):
    """The oracle (FP64 C restatement, OpenMP in the lattice step) on the
    rbc_in_pipe workload: same lattice, mesh, parameters."""
    import numpy as np
    from biblbm_amd import meshes, workloads
    from oracle import oracle_py as O
    R, length, u_c, tau = 32, 257, 0.02, 0.9
    mask, r2, c = workloads.pipe_mask(R, length)
    shape = mask.shape
    nu = (tau - 0.5) / 3.0
    g = 4.0 * nu * u_c / (R * R)
    prof = np.where(r2 <= R * R, u_c * (1.0 - r2 / (R * R)), 0.0)
    u = np.zeros(shape + (3,))
    u[..., 0] = prof[None]
    f = O.lbm_init_equilibrium(np.ones(shape), u)
    del u
    force = np.broadcast_to(np.array([g, 0, 0]), shape + (3,)).copy()
    v, t = meshes.biconcave(4, 12.0)
    vol = meshes.signed_volume(v, t)
    G = nu * (u_c / R) * 12.0 / 0.1
    k_bend = 2.0 * 0.02 * G * 144.0 / np.sqrt(3.0)
    rot = meshes.rotation_matrix([0, 1, 0], np.pi / 2)
    pos = v @ rot.T + [length / 2.0, c, c]
    part = O.Particle(v, t, 3, workloads.skalak_rbc_params(G, 10.0, 12.0, vol, k_bend), pos)
    phase = np.zeros(shape)
    done = 0
    t0 = time.perf_counter()
    while done < steps:
        f, force, _ = O.coupled_step(f, force, phase, mask, (1, 0, 0), 1 / tau, 1 / tau, (g, 0, 0), [part],
                                     "peskin")
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=round(mask.size * done / dt / 1e6, 3), unit="MLUPS", cores=cores, kind="port",
                sample="%d coupled steps of rbc_in_pipe (1 153 673 cells, 2562 vertices), FP64 oracle, "
                       "OpenMP in the lattice step" % done), part.pos


def suspension_leg(steps, warmup):
    import numpy as np
    from biblbm_amd import workloads
    out = {}
    u_ref = None
    for storage in ("f64", "f32", "f16s"):
        W = workloads.suspension(storage)
        dt = time_coupled(W, steps, warmup)
        _, u = W.lattice.moments()
        if storage == "f64":
            u_ref = u
            err = 0.0
        else:
            err = float(np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref))
        out[storage] = dict(mlups=round(W.cells * steps / dt / 1e6, 1), ms_per_step=round(dt / steps * 1e3, 4),
                            velocity_rel_l2_vs_f64=err)
        cells, nv, vf = W.cells, W.vertices, W.params["volume_fraction"]
        W.ibm.close()
        W.lattice.close()
        del W
    out["fp16s_over_fp32"] = round(out["f16s"]["mlups"] / out["f32"]["mlups"], 3)
    out["workload"] = ("suspension 512x128x128 channel, 400 cells (300x642, 80x2562, 20x162 vertices), "
                       "%.1f%% volume fraction, %d vertices, steps %d after %d warmup" %
                       (100 * vf, nv, steps, warmup))
    return out


def precision_leg(steps, checkpoints=(100, 200)):
    """384^3 periodic Taylor-Green box with 16 Skalak capsules (642
    vertices), stepped from the same start in FP64, FP32 and FP16-shifted
    storage; errors of the reduced forms against the FP64 run after each
    checkpoint step count and after `steps`: velocity field rel-L2, vertex
    positions rel-L2 and the largest per-capsule change of the Taylor
    deformation index.  Speed from the steps after the first checkpoint
    (graph replay, warm)."""
    import numpy as np
    from biblbm_amd import capi, workloads
    from biblbm_amd.ibm import taylor_deformation
    marks = sorted(set([c for c in checkpoints if c < steps] + [steps]))
    out = {}
    ref = {}
    for storage in ("f64", "f32", "f16s"):
        W = workloads.taylor_green(storage)
        row = {}
        done, t_timed, n_timed = 0, 0.0, 0
        for m in marks:
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


def ensemble_leg(steps, warmup, n_boxes=64):
    """64 independent shear boxes (sides from {32..64}^3, one 642-vertex cell
    each) advanced by one hipGraph launch per two steps, against the same
    boxes stepped one after another (each with its own captured graph)."""
    import numpy as np
    from biblbm_amd import workloads
    from biblbm_amd.ensemble import Ensemble, tune_collide
    from biblbm_amd.ibm import taylor_deformation
    ws = workloads.shear_ensemble("f64", n_boxes=n_boxes)
    cells = sum(w.cells for w in ws)
    tuned = tune_collide(ws[0].lattice)["best"]
    for w in ws:
        w.lattice.set_launch(tuned["block"], tuned["xcd_remap"])
    ens = Ensemble([(w.lattice, w.ibm) for w in ws])
    ens.steps(warmup)
    ms = ens.time_steps(steps)
    from biblbm_amd import capi
    for w in ws:
        w.ibm.coupled_steps(warmup, use_graph=True)
    capi.call("bl_sync")
    t0 = time.perf_counter()
    for w in ws:
        w.ibm.coupled_steps(steps, use_graph=True)
    capi.call("bl_sync")
    loop_s = time.perf_counter() - t0
    D = np.array([taylor_deformation(w.ibm.observables(0)["inertia"]) for w in ws])
    out = dict(mlups=round(cells * steps / (ms * 1e-3) / 1e6, 1), ms_per_step=round(ms / steps, 4),
               per_box_loop_mlups=round(cells * steps / loop_s / 1e6, 1),
               deformation_index=dict(min=float(D.min()), median=float(np.median(D)), max=float(D.max())),
               tuned_launch=tuned,
               workload="%d shear boxes, %d cells, sides %s, one 642-vertex biconcave cell each, FP64, "
                        "steps %d after %d warmup" % (n_boxes, cells, sorted({w.params["s"] for w in ws}), steps,
                                                      warmup))
    ens.close()
    for w in ws:
        w.ibm.close()
        w.lattice.close()
    return out


def progress(msg):
    """One line on stderr per finished leg (the JSON line stays alone on stdout)."""
    sys.stderr.write("bench: %s\n" % msg)
    sys.stderr.flush()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--storage", default="f64", choices=["f64", "f32", "f16s"])
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--cpu-budget", type=float, default=25.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        sys.exit("bench.py: this build runs on one GPU only (the path's multi-GPU sharding is a design, "
                 "not an implementation); run with --gpus 1")

    import numpy as np
    from biblbm_amd import capi, workloads

    capi.call("bl_init", 0)
    W = workloads.rbc_in_pipe(args.storage)
    dt = time_coupled(W, args.steps, args.warmup)
    value = W.cells * args.steps / dt / 1e6
    W.ibm.step()
    stages = {k: round(v, 5) for k, v in W.ibm.stage_times().items()}
    rl = roofline(W, args.storage)
    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": args.storage, "data": "synthetic",
        "config": {"workload": "rbc_in_pipe: pipe R=32 L=257 (257x67x67 cells, bounce-back), tau 0.9, body "
                               "force for u_c 0.02, one 2562-vertex Evans-Fung red cell R=12 (Skalak + bending "
                               "+ area + volume), Peskin kernel, hipGraph-replayed coupled steps",
                   "cells": W.cells, "fluid_cells": W.fluid_cells, "vertices": W.vertices,
                   "storage": args.storage, "parallelism": "single GPU"},
        "roofline": rl,
        "ibm_stage_ms": stages,
        "fluid_mlups": round(W.fluid_cells * args.steps / dt / 1e6, 2),
        "vertex_updates_per_s": round(W.vertices * args.steps / dt, 1),
    }
    W.ibm.close()
    W.lattice.close()
    if not args.no_extras:
        W32 = workloads.rbc_in_pipe("f32")
        dt32 = time_coupled(W32, args.steps, args.warmup)
        line["fp32"] = dict(value=round(W32.cells * args.steps / dt32 / 1e6, 2),
                            ms_per_step=round(dt32 / args.steps * 1e3, 5),
                            roofline=roofline(W32, "f32"))
        W32.ibm.close()
        W32.lattice.close()
        W2 = workloads.rbc_in_pipe("f64", second_order=True)
        dt2 = time_coupled(W2, args.steps // 2, args.warmup // 2)
        line["second_order"] = dict(value=round(W2.cells * (args.steps // 2) / dt2 / 1e6, 2),
                                    ms_per_step=round(dt2 / (args.steps // 2) * 1e3, 5),
                                    wall_nodes=W2.params["wall_nodes"], fluid_cells=W2.fluid_cells)
        W2.ibm.close()
        W2.guo.close()
        W2.lattice.close()
        progress("rbc_in_pipe f64 / f32 / second order done")
        line["suspension"] = suspension_leg(max(50, args.steps // 10), 20)
        progress("suspension done")
        line["interactions"] = interactions_leg(max(50, args.steps // 10), 20,
                                                line["suspension"]["f32"]["ms_per_step"])
        progress("interactions done")
        line["ensemble"] = ensemble_leg(max(50, args.steps // 10), 10)
        progress("ensemble done")
        line["precision"] = precision_leg(max(300, args.steps // 4))
        progress("precision done")
    if not args.no_cpu_baseline:
        cb, cpu_pos = cpu_baseline(args.cpu_steps, args.cpu_budget)
        # the same number of coupled steps on the GPU from the same state: parity check
        n = int(cb["sample"].split()[0])
        G = workloads.rbc_in_pipe("f64")
        G.ibm.coupled_steps(n, use_graph=False)
        gpos = G.ibm.positions()
        cb["gpu_vs_cpu_position_rel_l2"] = float(np.linalg.norm(gpos - cpu_pos) / np.linalg.norm(cpu_pos))
        line["cpu_baseline"] = cb
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
